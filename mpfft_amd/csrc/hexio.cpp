// Exact text form of SoA-limb data: hexadecimal floating point in the
// format MPFR's printf prints for "%Ra" (the format of the reference's
// mpfr_fprintf calls with the 'a' conversion), so fixtures and dumps can be
// read and diffed without loss.  Host code only.
//
// Format of a nonzero value x = 0.1b...b * 2^E (MPFR convention):
//   [-]0x<d>[.<hex digits>]p<+|-><decimal exponent>
// where the printed binary exponent is e = 4*floor((E-1)/4), the leading
// digit d = x / 2^e lies in [1, 16), and trailing zero digits are dropped
// (so "0x1p+0" for 1, "0xcp-4" for 0.75).  Zero prints as "0x0p+0" or
// "-0x0p+0".
#include "../../include/mpfft.h"

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr uint32_t kHdrZeroU = 0x80000000u;

long floor_div4(long v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }

// one real: header h, limbs lim[j * M + i], j = 0 most significant
std::string format_one(uint32_t h, const uint32_t *lim, size_t M, size_t i, long L)
{
    const bool neg = h & 1u;
    std::string s = neg ? "-0x" : "0x";
    if ((h & ~1u) == kHdrZeroU) return s + "0p+0";
    const long E = (int32_t)h >> 1;
    const long e = 4 * floor_div4(E - 1);
    const int lead = (int)(E - e);                    // bits in the leading digit, 1..4
    // significand bits, MSB first
    std::vector<uint8_t> bits((size_t)L * 32);
    for (long j = 0; j < L; j++) {
        const uint32_t w = lim[(size_t)j * M + i];
        for (int b = 0; b < 32; b++) bits[(size_t)j * 32 + b] = (w >> (31 - b)) & 1u;
    }
    int d = 0;
    for (int b = 0; b < lead; b++) d = 2 * d + bits[b];
    static const char hx[] = "0123456789abcdef";
    s += hx[d];
    std::string frac;
    for (size_t b = lead; b < bits.size(); b += 4) {
        int v = 0;
        for (size_t t = 0; t < 4; t++) v = 2 * v + (b + t < bits.size() ? bits[b + t] : 0);
        frac += hx[v];
    }
    while (!frac.empty() && frac.back() == '0') frac.pop_back();
    if (!frac.empty()) s += "." + frac;
    char ex[32];
    snprintf(ex, sizeof ex, "p%+ld", e);
    return s + ex;
}

int hexval(int c)
{
    if (c >= '0' && c <= '9') return c - '0';
    c = tolower(c);
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    return -1;
}

// parse one hex float at *p (advanced past it); round to prec bits (nearest,
// ties to even) into header and L limbs
int parse_one(const char *&p, long prec, long L, uint32_t *h, std::vector<uint32_t> &limbs)
{
    while (isspace((unsigned char)*p)) p++;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    if (p[0] != '0' || (p[1] != 'x' && p[1] != 'X')) return MPFFT_ERR_VALUE;
    p += 2;
    std::vector<uint8_t> bits;                    // significand bits, MSB first
    long point = -1;                              // bits before the radix point
    int ndig = 0;
    for (;; p++) {
        if (*p == '.') {
            if (point >= 0) return MPFFT_ERR_VALUE;
            point = (long)bits.size();
            continue;
        }
        const int v = hexval((unsigned char)*p);
        if (v < 0) break;
        for (int b = 3; b >= 0; b--) bits.push_back((v >> b) & 1);
        ndig++;
    }
    if (ndig == 0) return MPFFT_ERR_VALUE;
    if (point < 0) point = (long)bits.size();
    long pexp = 0;
    if (*p == 'p' || *p == 'P') {
        p++;
        bool eneg = false;
        if (*p == '+' || *p == '-') eneg = *p++ == '-';
        if (!isdigit((unsigned char)*p)) return MPFFT_ERR_VALUE;
        while (isdigit((unsigned char)*p)) {
            pexp = 10 * pexp + (*p++ - '0');
            if (pexp > (1L << 40)) return MPFFT_ERR_VALUE;
        }
        if (eneg) pexp = -pexp;
    }
    if (*p && !isspace((unsigned char)*p)) return MPFFT_ERR_VALUE;
    size_t first = 0;
    while (first < bits.size() && !bits[first]) first++;
    std::fill(limbs.begin(), limbs.end(), 0u);
    if (first == bits.size()) {                   // zero
        *h = kHdrZeroU | (neg ? 1u : 0u);
        return MPFFT_OK;
    }
    // value = 0.b[first] b[first+1] ... * 2^E
    long E = point - (long)first + pexp;
    std::vector<uint8_t> m(bits.begin() + (long)first, bits.end());
    if ((long)m.size() > prec) {
        const bool guard = m[prec];
        bool sticky = false;
        for (size_t t = (size_t)prec + 1; t < m.size(); t++) sticky |= m[t];
        m.resize(prec);
        if (guard && (sticky || m[prec - 1])) {
            long t = prec - 1;
            while (t >= 0 && m[t]) m[t--] = 0;
            if (t < 0) { m.assign(prec, 0); m[0] = 1; E += 1; }
            else m[t] = 1;
        }
    }
    for (size_t t = 0; t < m.size() && t < (size_t)L * 32; t++)
        if (m[t]) limbs[t / 32] |= 1u << (31 - t % 32);
    if (E >= (1L << 29) || E <= -(1L << 29)) return MPFFT_ERR_VALUE;
    *h = ((uint32_t)E << 1) | (neg ? 1u : 0u);
    return MPFFT_OK;
}

}  // namespace

extern "C" {

int mpfft_limbs_to_hex(unsigned long count, long limbs, const uint32_t *h_in, char *buf,
                       size_t bufsize, size_t *needed)
{
    if (!h_in || limbs < 1 || (!buf && bufsize)) return MPFFT_ERR_ARG;
    std::string out;
    const uint32_t *hdr = h_in, *lim = h_in + count;
    for (unsigned long i = 0; i < count; i++) {
        out += format_one(hdr[i], lim, count, i, limbs);
        out += '\n';
    }
    if (needed) *needed = out.size() + 1;
    if (out.size() + 1 > bufsize) return MPFFT_ERR_ARG;
    memcpy(buf, out.c_str(), out.size() + 1);
    return MPFFT_OK;
}

int mpfft_hex_to_limbs(unsigned long count, long limbs, long prec, const char *text,
                       uint32_t *h_out)
{
    if (!text || !h_out || limbs < 1 || prec < 1 || prec > 32 * limbs) return MPFFT_ERR_ARG;
    std::vector<uint32_t> lim((size_t)limbs);
    const char *p = text;
    for (unsigned long i = 0; i < count; i++) {
        uint32_t h;
        const int rc = parse_one(p, prec, limbs, &h, lim);
        if (rc) return rc;
        h_out[i] = h;
        for (long j = 0; j < limbs; j++) h_out[count + (size_t)j * count + i] = lim[(size_t)j];
    }
    return MPFFT_OK;
}

}  // extern "C"
