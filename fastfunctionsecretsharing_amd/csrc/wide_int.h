// Unsigned integers of W 64-bit limbs (little-endian limb order), arithmetic
// modulo 2^(64 W): the group elements of 16-, 24- and 32-byte elementSize
// (the reference's wideint::uint<2>, <3>, <4>, src/impl/def.h:37-44).  In
// memory a value is its W limbs, least significant first, so its bytes are
// the little-endian bytes of the integer, as for the built-in widths.
//
// Conversions follow C's unsigned rules: from a signed integer by sign
// extension (WideInt<W>(-1) is all ones), to a narrower unsigned integer by
// truncation.  Shifts by at least 64 W bits give 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fss {

template <int W>
struct WideInt
{
    static_assert(W >= 2 && W <= 4, "16- to 32-byte elements");
    uint64_t l[W];

    WideInt() = default;

    template <typename I, typename = std::enable_if_t<std::is_integral<I>::value>>
    __host__ __device__ constexpr WideInt(I v) : l{}
    {
        l[0]             = (uint64_t)v;
        const uint64_t h = (std::is_signed<I>::value && v < 0) ? ~0ull : 0ull;
        for (int i = 1; i < W; i++) l[i] = h;
    }

    template <typename I, typename = std::enable_if_t<std::is_integral<I>::value>>
    __host__ __device__ explicit constexpr operator I() const
    {
        return (I)l[0];
    }

    __host__ __device__ explicit constexpr operator bool() const
    {
        uint64_t o = 0;
        for (int i = 0; i < W; i++) o |= l[i];
        return o != 0;
    }

    // ---- bitwise ----
    __host__ __device__ friend constexpr WideInt operator~(WideInt a)
    {
        for (int i = 0; i < W; i++) a.l[i] = ~a.l[i];
        return a;
    }
    __host__ __device__ friend constexpr WideInt operator&(WideInt a, const WideInt &b)
    {
        for (int i = 0; i < W; i++) a.l[i] &= b.l[i];
        return a;
    }
    __host__ __device__ friend constexpr WideInt operator|(WideInt a, const WideInt &b)
    {
        for (int i = 0; i < W; i++) a.l[i] |= b.l[i];
        return a;
    }
    __host__ __device__ friend constexpr WideInt operator^(WideInt a, const WideInt &b)
    {
        for (int i = 0; i < W; i++) a.l[i] ^= b.l[i];
        return a;
    }
    __host__ __device__ friend constexpr WideInt operator<<(const WideInt &a, int s)
    {
        WideInt r(0);
        if (s < 0 || s >= 64 * W) return r;
        const int q = s >> 6, b = s & 63;
        for (int i = W - 1; i >= q; i--)
        {
            uint64_t v = a.l[i - q] << b;
            if (b && i - q - 1 >= 0) v |= a.l[i - q - 1] >> (64 - b);
            r.l[i] = v;
        }
        return r;
    }
    __host__ __device__ friend constexpr WideInt operator>>(const WideInt &a, int s)
    {
        WideInt r(0);
        if (s < 0 || s >= 64 * W) return r;
        const int q = s >> 6, b = s & 63;
        for (int i = 0; i + q < W; i++)
        {
            uint64_t v = a.l[i + q] >> b;
            if (b && i + q + 1 < W) v |= a.l[i + q + 1] << (64 - b);
            r.l[i] = v;
        }
        return r;
    }

    // ---- arithmetic modulo 2^(64 W) ----
    __host__ __device__ friend constexpr WideInt operator+(const WideInt &a, const WideInt &b)
    {
        WideInt  r(0);
        uint64_t carry = 0;
        for (int i = 0; i < W; i++)
        {
            const uint64_t s = a.l[i] + b.l[i];
            const uint64_t c = s < a.l[i];
            r.l[i]           = s + carry;
            carry            = c | (r.l[i] < s);
        }
        return r;
    }
    __host__ __device__ friend constexpr WideInt operator-(const WideInt &a, const WideInt &b)
    {
        WideInt  r(0);
        uint64_t borrow = 0;
        for (int i = 0; i < W; i++)
        {
            const uint64_t d  = a.l[i] - b.l[i];
            const uint64_t b1 = a.l[i] < b.l[i];
            r.l[i]            = d - borrow;
            borrow            = b1 | (d < borrow);
        }
        return r;
    }
    __host__ __device__ friend constexpr WideInt operator-(const WideInt &a) { return WideInt(0) - a; }
    // low 64 W bits of the product (schoolbook on 64-bit limbs)
    __host__ __device__ friend constexpr WideInt operator*(const WideInt &a, const WideInt &b)
    {
        WideInt r(0);
        for (int i = 0; i < W; i++)
        {
            uint64_t carry = 0;
            for (int j = 0; i + j < W; j++)
            {
                const unsigned __int128 p =
                    (unsigned __int128)a.l[i] * b.l[j] + r.l[i + j] + (unsigned __int128)carry;
                r.l[i + j] = (uint64_t)p;
                carry      = (uint64_t)(p >> 64);
            }
        }
        return r;
    }

    // ---- comparisons (unsigned) ----
    __host__ __device__ friend constexpr bool operator==(const WideInt &a, const WideInt &b)
    {
        uint64_t d = 0;
        for (int i = 0; i < W; i++) d |= a.l[i] ^ b.l[i];
        return d == 0;
    }
    __host__ __device__ friend constexpr bool operator!=(const WideInt &a, const WideInt &b) { return !(a == b); }
    __host__ __device__ friend constexpr bool operator<(const WideInt &a, const WideInt &b)
    {
        for (int i = W - 1; i >= 0; i--)
            if (a.l[i] != b.l[i]) return a.l[i] < b.l[i];
        return false;
    }
    __host__ __device__ friend constexpr bool operator>(const WideInt &a, const WideInt &b) { return b < a; }
    __host__ __device__ friend constexpr bool operator<=(const WideInt &a, const WideInt &b) { return !(b < a); }
    __host__ __device__ friend constexpr bool operator>=(const WideInt &a, const WideInt &b) { return !(a < b); }

    // byte k (0 = least significant)
    __host__ __device__ constexpr uint32_t byte(int k) const { return (uint32_t)(l[k >> 3] >> (8 * (k & 7))) & 0xffu; }
};

// x mod 2^bw (bw >= 64 W: x itself)
template <int W>
__host__ __device__ constexpr WideInt<W> wide_mod_bits(const WideInt<W> &x, int bw)
{
    if (bw >= 64 * W) return x;
    return x & ((WideInt<W>(1) << bw) - WideInt<W>(1));
}

} // namespace fss
