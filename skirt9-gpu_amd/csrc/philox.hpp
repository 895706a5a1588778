// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw, SC'11), usable on the
// host and on the device.  A stream is identified by (seed, history, tag); draw i uses block
// i/2 of the stream and the word pair (2*(i%2), 2*(i%2)+1), giving a 52-bit integer a and the
// uniform deviate (a + 0.5) * 2^-52 on the open interval (0,1) (SKIRT's Random excludes both
// ends too, Random.cpp:24-27).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace skg
{
    __host__ __device__ __forceinline__ void philox4x32_10(uint32_t& x0, uint32_t& x1, uint32_t& x2, uint32_t& x3,
                                                            uint32_t k0, uint32_t k1)
    {
#pragma unroll
        for (int r = 0; r < 10; ++r)
        {
            if (r)
            {
                k0 += 0x9E3779B9u;
                k1 += 0xBB67AE85u;
            }
            uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * x0;
            uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * x2;
            uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ x1 ^ k0;
            uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ x3 ^ k1;
            x1 = static_cast<uint32_t>(p1);
            x3 = static_cast<uint32_t>(p0);
            x0 = n0;
            x2 = n2;
        }
    }

    struct Philox
    {
        uint32_t k0, k1, h0, h1, tag;
        uint32_t draw;
        uint32_t block;
        uint32_t c[4];

        __host__ __device__ __forceinline__ void init(uint64_t seed, uint64_t history, uint32_t tg = 0)
        {
            k0 = static_cast<uint32_t>(seed);
            k1 = static_cast<uint32_t>(seed >> 32);
            h0 = static_cast<uint32_t>(history);
            h1 = static_cast<uint32_t>(history >> 32);
            tag = tg;
            draw = 0;
            block = 0xFFFFFFFFu;
        }

        // continue the same history on the stream of another tag, from its first draw
        __host__ __device__ __forceinline__ void retag(uint32_t tg)
        {
            tag = tg;
            draw = 0;
            block = 0xFFFFFFFFu;
        }

        __host__ __device__ __forceinline__ void gen(uint32_t b)
        {
            uint32_t x0 = h0, x1 = h1, x2 = b, x3 = tag;
            philox4x32_10(x0, x1, x2, x3, k0, k1);
            c[0] = x0;
            c[1] = x1;
            c[2] = x2;
            c[3] = x3;
            block = b;
        }

        __host__ __device__ __forceinline__ double uniform()
        {
            uint32_t b = draw >> 1;
            if (b != block) gen(b);
            uint32_t hi = (draw & 1) ? c[2] : c[0];
            uint32_t lo = (draw & 1) ? c[3] : c[1];
            ++draw;
            uint64_t a = (static_cast<uint64_t>(hi) << 20) ^ (lo >> 12);
            return (static_cast<double>(a) + 0.5) * (1.0 / 4503599627370496.0);
        }
    };
}
