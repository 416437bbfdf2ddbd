// crc_math.h -- CRC-32/ISO-HDLC arithmetic in GF(2)[x] / P, reflected bit order
// (poly 0xEDB88320), usable from host and device code.
//
// The device inflate computes the CRC of a member as independent lane segments
// and folds them with these identities (all on the zero-initialised register
// "raw" CRC, which is linear):
//   raw(A || B)       = mulmod(raw(A), xpow8(|B|)) ^ raw(B)
//   raw(0^k || A)     = raw(A)
//   crc32(M)          = raw(M) ^ mulmod(0xFFFFFFFF, xpow8(|M|)) ^ 0xFFFFFFFF
// where xpow8(n) = x^(8n) mod P. In this representation x^0 is 0x80000000.
#ifndef FRHIP_CRC_MATH_H
#define FRHIP_CRC_MATH_H
#include <stdint.h>

#ifndef FR_HD
#if defined(__HIPCC__)
#define FR_HD __host__ __device__
#else
#define FR_HD
#endif
#endif

#define FR_CRC_POLY 0xEDB88320u

// a * b mod P, fixed 32-step shift-and-add (no data-dependent branch).
static inline FR_HD uint32_t fr_mulmod(uint32_t a, uint32_t b)
{
    uint32_t p = 0;
    for (int k = 0; k < 32; ++k) {
        p ^= b & (0u - ((a >> (31 - k)) & 1u));
        b = (b >> 1) ^ (FR_CRC_POLY & (0u - (b & 1u)));
    }
    return p;
}

// x^(2^k) mod P for k = 0..31, filled by fr_x2n_table.
static inline FR_HD void fr_x2n_table(uint32_t *t)
{
    uint32_t p = 1u << 30; // x^1
    t[0] = p;
    for (int k = 1; k < 32; ++k) t[k] = p = fr_mulmod(p, p);
}

// x^(8n) mod P from the x^(2^k) table.
static inline FR_HD uint32_t fr_xpow8(const uint32_t *x2n, uint64_t n)
{
    uint32_t p = 1u << 31; // x^0
    int k = 3;
    while (n) {
        if (n & 1) p = fr_mulmod(x2n[k & 31], p);
        n >>= 1;
        ++k;
    }
    return p;
}

// Byte-wise table for the reflected polynomial.
static inline FR_HD void fr_crc_table(uint32_t *t)
{
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (FR_CRC_POLY & (0u - (c & 1u)));
        t[i] = c;
    }
}

#endif
