// spd_decimal.hpp -- text-to-number conversion for the .tns reader, the same
// on the host and on the device.
//
// The reference converts each field with strtoull (indices) and strtod
// (the value) (src/io.c:273-348). These functions accept the common shapes
// of those fields and return exactly what strtoull / strtod return for them;
// for anything else they decline (return false) and the caller converts the
// whole line with the C library:
//   index: [blanks] [+] 1..19 digits, followed by a blank or the line's end
//   value: [blanks] [+|-] (digits [. [digits]] | . digits) [(e|E) [+|-] digits],
//          followed by a blank or the line's end
// with at most 19 significant digits and a result that is zero or a normal
// double (no overflow, no subnormal). Blanks are the C locale's isspace
// characters other than '\n' (the line ends there).
//
// Decimal to binary: the Eisel-Lemire method (D. Lemire, "Number Parsing at
// a Gigabyte per Second", Software: Practice and Experience 51(8), 2021):
// the 19-digit significand w, shifted to bit 63, times a 128-bit
// approximation of 5^q (spd_pow5.inc); the top bits give the rounded
// mantissa. When the product cannot decide the rounding (the truncated bits
// are all ones and the exponent lies outside the range where the product is
// exact), the conversion declines.
#pragma once
#include <stdint.h>

#define SPD_POW5_MIN (-342)
#define SPD_POW5_MAX 308

__host__ __device__ static inline bool spd_blank(char c)
{
  return c == ' ' || c == '\t' || c == '\v' || c == '\f' || c == '\r';
}

// Index field at [*pp, e): on success *pp is past the digits.
__host__ __device__ static inline bool spd_parse_index(char const ** pp, char const * e,
    uint64_t * out)
{
  char const * p = *pp;
  while(p < e && spd_blank(*p)) ++p;
  if(p < e && *p == '+') ++p;
  uint64_t v = 0;
  int nd = 0;
  while(p < e && *p >= '0' && *p <= '9') {
    if(++nd > 19) return false;
    v = v * 10u + (uint64_t) (*p - '0');
    ++p;
  }
  if(nd == 0) return false;
  if(p < e && !spd_blank(*p)) return false;
  *pp = p;
  *out = v;
  return true;
}

__host__ __device__ static inline void spd_mul128(uint64_t a, uint64_t b, uint64_t * hi,
    uint64_t * lo)
{
  unsigned __int128 const r = (unsigned __int128) a * b;
  *hi = (uint64_t) (r >> 64);
  *lo = (uint64_t) r;
}

// w * 10^q rounded to nearest even, w != 0; false when undecided or out of
// the normal range.
__host__ __device__ static inline bool spd_eisel_lemire(uint64_t w, int q,
    uint64_t const * pow5, uint64_t * bits)
{
  if(q < SPD_POW5_MIN || q > SPD_POW5_MAX) return false;
  int const lz = __builtin_clzll(w);
  w <<= lz;
  int const idx = 2 * (q - SPD_POW5_MIN);
  uint64_t hi, lo;
  spd_mul128(w, pow5[idx], &hi, &lo);
  uint64_t const mask = 0x1FFull;       // 64 - 52 - 3 = 9 bits below the mantissa
  if((hi & mask) == mask) {
    uint64_t hi2, lo2;
    spd_mul128(w, pow5[idx + 1], &hi2, &lo2);
    lo += hi2;
    if(hi2 > lo) ++hi;
  }
  // the truncated product may be one unit short of a carry into the
  // mantissa; exact only for 5^q that fits the 128-bit entry
  if(lo == 0xFFFFFFFFFFFFFFFFull && (q < -27 || q > 55)) return false;
  int const upper = (int) (hi >> 63);
  int const shift = upper + 9;
  uint64_t m = hi >> shift;
  // binary exponent: floor(q * log2(10)) + 63 - lz + upper + bias
  int p2 = (int) ((((int64_t) (152170 + 65536) * q) >> 16) + 63) + upper - lz + 1023;
  if(p2 <= 0) return false;             // subnormal or zero: leave it to strtod
  // exactly halfway between two doubles: round to even
  if(lo <= 1 && q >= -4 && q <= 23 && (m & 3) == 1 && (m << shift) == hi) {
    m &= ~(uint64_t) 1;
  }
  m += m & 1;
  m >>= 1;
  if(m >= (2ull << 52)) {
    m = 1ull << 52;
    ++p2;
  }
  m &= ~(1ull << 52);
  if(p2 >= 0x7FF) return false;         // overflow: leave it to strtod
  *bits = m | ((uint64_t) p2 << 52);
  return true;
}

// Value field at [*pp, e).
__host__ __device__ static inline bool spd_parse_value(char const ** pp, char const * e,
    uint64_t const * pow5, double * out)
{
  char const * p = *pp;
  while(p < e && spd_blank(*p)) ++p;
  bool neg = false;
  if(p < e && (*p == '+' || *p == '-')) {
    neg = (*p == '-');
    ++p;
  }
  uint64_t w = 0;
  int nsig = 0;            // significant digits taken into w
  int ndig = 0;            // digits seen
  int q = 0;               // decimal exponent of w
  while(p < e && *p >= '0' && *p <= '9') {
    ++ndig;
    if(w != 0 || *p != '0') {
      if(++nsig > 19) return false;
      w = w * 10u + (uint64_t) (*p - '0');
    }
    ++p;
  }
  if(p < e && *p == '.') {
    ++p;
    while(p < e && *p >= '0' && *p <= '9') {
      ++ndig;
      if(w != 0 || *p != '0') {
        if(++nsig > 19) return false;
        w = w * 10u + (uint64_t) (*p - '0');
      }
      --q;
      ++p;
    }
  }
  if(ndig == 0) return false;
  if(p < e && (*p == 'e' || *p == 'E')) {
    char const * x = p + 1;
    bool eneg = false;
    if(x < e && (*x == '+' || *x == '-')) {
      eneg = (*x == '-');
      ++x;
    }
    if(x < e && *x >= '0' && *x <= '9') {
      int ev = 0;
      while(x < e && *x >= '0' && *x <= '9') {
        if(ev > 100000) return false;
        ev = ev * 10 + (*x - '0');
        ++x;
      }
      q += eneg ? -ev : ev;
      p = x;
    }
  }
  // strtod stops here and the rest of the line is not read; but a field
  // that goes on with other characters may be one that strtod reads further
  // ("0x1p3", "1e5" followed by letters), so only a blank or the line's end
  // is accepted after it
  if(p < e && !spd_blank(*p)) return false;
  uint64_t bits = 0;
  if(w != 0 && !spd_eisel_lemire(w, q, pow5, &bits)) return false;
  if(neg) bits |= 1ull << 63;
  double v;
  __builtin_memcpy(&v, &bits, sizeof(v));
  *out = v;
  *pp = p;
  return true;
}
