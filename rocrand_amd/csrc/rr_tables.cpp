// rr_tables.cpp -- jump-ahead tables of the stateful engines, computed on the host
// from the one-step recurrences and uploaded once per device.
//
// The reference ships these as generated headers (rocrand_mrg32k3a_precomputed.h,
// rocrand_mrg31k3p_precomputed.h, rocrand_xorwow_precomputed.h,
// rocrand_lfsr113_precomputed.h, written by tools/*_precomputed_generator.cpp).
// They are pure matrix powers, so they are derived here instead of carried:
//   MRG32k3a: A^(2^i) and A^(2^(76+i))       (subsequence = 2^76 steps)
//   MRG31k3p: A^(2^i) and A^(2^(72+i))       (subsequence = 2^72 steps)
//   both    : radix-512/256 subsequence tables (engine e < 2^17 in two products)
//   XORWOW  : M^(4^k) and M^(2^67 * 4^k)     (GF(2), 160x160)
//   LFSR113 : M^(4^k) and M^(2^55 * 4^k)     (GF(2), 128x128)
#include "rr_launch.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string.h>
#include <vector>

namespace rr
{
// The device copies live in module-scope device arrays (one instance per GPU), filled by
// a copy on a device's first use: no device allocation is ever made for them.
__device__ uint32_t d_mrg32_step[64 * 18], d_mrg32_sub[64 * 18], d_mrg31_step[64 * 18], d_mrg31_sub[64 * 18];
__device__ uint32_t d_mrg32_subr[(MRG_RADIX_LO + MRG_RADIX_HI) * MRG_RADIX_WORDS];
__device__ uint32_t d_mrg31_subr[(MRG_RADIX_LO + MRG_RADIX_HI) * MRG_RADIX_WORDS];
__device__ uint32_t d_xorwow_jump[32 * 800], d_xorwow_seq[32 * 800], d_lfsr_jump[32 * 512], d_lfsr_seq[32 * 512];

namespace
{

// ---- mod-m 3x3 matrices (row-major, entries < m)
void mat3_mul(const uint64_t* A, const uint64_t* B, uint64_t* C, uint64_t m)
{
    uint64_t T[9];
    for(int r = 0; r < 3; r++)
        for(int c = 0; c < 3; c++)
        {
            uint64_t acc = 0;
            for(int k = 0; k < 3; k++)
                acc += (A[3 * r + k] * B[3 * k + c]) % m;
            T[3 * r + c] = acc % m;
        }
    memcpy(C, T, sizeof(T));
}

// table[i][comp][9] = X_comp^(2^(first + i)), i = 0..63
void mrg_table(const uint64_t A1[9], const uint64_t A2[9], uint64_t m1, uint64_t m2, int first, uint32_t* out)
{
    const uint64_t* A[2] = {A1, A2};
    const uint64_t  m[2] = {m1, m2};
    for(int c = 0; c < 2; c++)
    {
        uint64_t X[9];
        memcpy(X, A[c], sizeof(X));
        for(int i = 0; i < first; i++)
            mat3_mul(X, X, X, m[c]);
        for(int i = 0; i < 64; i++)
        {
            for(int j = 0; j < 9; j++)
                out[(i * 2 + c) * 9 + j] = (uint32_t)X[j];
            mat3_mul(X, X, X, m[c]);
        }
    }
}

// Radix tables of the subsequence jump: entry d (d < 512) holds X^d and entry 512 + d
// (d < 256) holds X^(512 d), X = A^(2^first) per component, so an engine index below
// 2^17 costs two table products.  Entry layout: component 0 (9 words), component 1
// (9 words), 2 words of padding (16-byte aligned entries).
void mrg_radix_table(const uint32_t* pow2, uint64_t m1, uint64_t m2, uint32_t* out)
{
    const uint64_t m[2] = {m1, m2};
    for(int c = 0; c < 2; c++)
        for(int level = 0; level < 2; level++)
        {
            uint64_t X[9], P[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            const int i = level == 0 ? 0 : MRG_RADIX_LO_BITS; // X^(2^i) from the power table
            for(int j = 0; j < 9; j++)
                X[j] = pow2[(i * 2 + c) * 9 + j];
            const int first = level == 0 ? 0 : MRG_RADIX_LO;
            const int count = level == 0 ? MRG_RADIX_LO : MRG_RADIX_HI;
            for(int d = 0; d < count; d++)
            {
                for(int j = 0; j < 9; j++)
                    out[(first + d) * MRG_RADIX_WORDS + c * 9 + j] = (uint32_t)P[j];
                mat3_mul(P, X, P, m[c]);
            }
        }
}

// ---- GF(2) matrices: row (word i, bit j) holds the N-word image of that input bit.
template<int N>
void gf2_apply(const uint32_t* m, uint32_t* v)
{
    uint32_t r[N] = {};
    for(int i = 0; i < N; i++)
        for(int j = 0; j < 32; j++)
            if((v[i] >> j) & 1u)
                for(int k = 0; k < N; k++)
                    r[k] ^= m[(i * 32 + j) * N + k];
    memcpy(v, r, sizeof(r));
}
template<int N>
void gf2_mul(const uint32_t* A, const uint32_t* B, uint32_t* C) // apply A then B
{
    std::vector<uint32_t> T(32 * N * N);
    for(int row = 0; row < 32 * N; row++)
    {
        uint32_t v[N];
        memcpy(v, A + row * N, sizeof(v));
        gf2_apply<N>(B, v);
        memcpy(&T[row * N], v, sizeof(v));
    }
    memcpy(C, T.data(), T.size() * sizeof(uint32_t));
}
template<int N>
void gf2_tables(void (*step)(uint32_t*), int seq_log2, uint32_t* jump, uint32_t* seq)
{
    const int             SZ = 32 * N * N;
    std::vector<uint32_t> one(SZ), a(SZ);
    for(int i = 0; i < N; i++)
        for(int j = 0; j < 32; j++)
        {
            uint32_t v[N] = {};
            v[i]          = 1u << j;
            step(v);
            memcpy(&one[(i * 32 + j) * N], v, sizeof(v));
        }
    a = one;
    for(int k = 0; k < 32; k++)
    {
        memcpy(jump + k * SZ, a.data(), SZ * 4);
        gf2_mul<N>(a.data(), a.data(), a.data());
        gf2_mul<N>(a.data(), a.data(), a.data());
    }
    a = one;
    for(int s = 0; s < seq_log2; s++)
        gf2_mul<N>(a.data(), a.data(), a.data());
    for(int k = 0; k < 32; k++)
    {
        memcpy(seq + k * SZ, a.data(), SZ * 4);
        gf2_mul<N>(a.data(), a.data(), a.data());
        gf2_mul<N>(a.data(), a.data(), a.data());
    }
}

void xorwow_x_step(uint32_t* x)
{
    xorwow_state s{0, {x[0], x[1], x[2], x[3], x[4]}};
    xorwow_next(s);
    memcpy(x, s.x, 5 * 4);
}
void lfsr_step(uint32_t* z) { lfsr113_step(z); }

struct host_tables
{
    std::vector<uint32_t> mrg32_step, mrg32_sub, mrg31_step, mrg31_sub, mrg32_subr, mrg31_subr;
    std::vector<uint32_t> xorwow_jump, xorwow_seq, lfsr_jump, lfsr_seq;
    host_tables()
    {
        mrg32_step.resize(64 * 18);
        mrg32_sub.resize(64 * 18);
        mrg31_step.resize(64 * 18);
        mrg31_sub.resize(64 * 18);
        {
            // MRG32k3a one-step matrices, state (oldest .. newest), rocrand_mrg32k3a.h:160-193
            const uint64_t A1[9] = {0, 1, 0, 0, 0, 1, MRG32_M1 - MRG32_A13N, MRG32_A12, 0};
            const uint64_t A2[9] = {0, 1, 0, 0, 0, 1, MRG32_M2 - MRG32_A23N, 0, MRG32_A21};
            mrg_table(A1, A2, MRG32_M1, MRG32_M2, 0, mrg32_step.data());
            mrg_table(A1, A2, MRG32_M1, MRG32_M2, 76, mrg32_sub.data());
        }
        {
            // MRG31k3p one-step matrices, state (newest .. oldest), rocrand_mrg31k3p.h next()
            const uint64_t A1[9] = {0, 1ull << 22, (1ull << 7) + 1, 1, 0, 0, 0, 1, 0};
            const uint64_t A2[9] = {1ull << 15, 0, (1ull << 15) + 1, 1, 0, 0, 0, 1, 0};
            mrg_table(A1, A2, MRG31_M1, MRG31_M2, 0, mrg31_step.data());
            mrg_table(A1, A2, MRG31_M1, MRG31_M2, 72, mrg31_sub.data());
        }
        mrg32_subr.assign((MRG_RADIX_LO + MRG_RADIX_HI) * MRG_RADIX_WORDS, 0u);
        mrg31_subr.assign((MRG_RADIX_LO + MRG_RADIX_HI) * MRG_RADIX_WORDS, 0u);
        mrg_radix_table(mrg32_sub.data(), MRG32_M1, MRG32_M2, mrg32_subr.data());
        mrg_radix_table(mrg31_sub.data(), MRG31_M1, MRG31_M2, mrg31_subr.data());
        xorwow_jump.resize(32 * 800);
        xorwow_seq.resize(32 * 800);
        gf2_tables<5>(xorwow_x_step, 67, xorwow_jump.data(), xorwow_seq.data());
        lfsr_jump.resize(32 * 512);
        lfsr_seq.resize(32 * 512);
        gf2_tables<4>(lfsr_step, 55, lfsr_jump.data(), lfsr_seq.data());
    }
};

const host_tables& host()
{
    static host_tables t;
    return t;
}

struct device_cache
{
    std::mutex    mu;
    bool          ready[64] = {};
    device_tables t[64];
};
device_cache& cache()
{
    static device_cache c;
    return c;
}

// Copies a host table into the current device's instance of `symbol`.
hipError_t upload(const std::vector<uint32_t>& v, const void* symbol, size_t symbol_bytes, const uint32_t** out)
{
    if(v.size() * 4 != symbol_bytes)
        return hipErrorInvalidValue;
    void*      p   = nullptr;
    hipError_t err = hipGetSymbolAddress(&p, symbol);
    if(err != hipSuccess)
        return err;
    err  = hipMemcpy(p, v.data(), symbol_bytes, hipMemcpyHostToDevice);
    *out = static_cast<const uint32_t*>(p);
    return err;
}

} // namespace

hipError_t get_device_tables(device_tables& t)
{
    int        dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if(err != hipSuccess)
        return err;
    if(dev < 0 || dev >= 64)
        return hipErrorInvalidDevice;
    device_cache&               c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    if(!c.ready[dev])
    {
        const host_tables& h = host();
        device_tables      d{};
#define RR_UPLOAD(name) \
    if((err = upload(h.name, HIP_SYMBOL(d_##name), sizeof(d_##name), &d.name)) != hipSuccess) return err;
        RR_UPLOAD(mrg32_step)
        RR_UPLOAD(mrg32_sub)
        RR_UPLOAD(mrg31_step)
        RR_UPLOAD(mrg31_sub)
        RR_UPLOAD(mrg32_subr)
        RR_UPLOAD(mrg31_subr)
        RR_UPLOAD(xorwow_jump)
        RR_UPLOAD(xorwow_seq)
        RR_UPLOAD(lfsr_jump)
        RR_UPLOAD(lfsr_seq)
#undef RR_UPLOAD
        c.t[dev]     = d;
        c.ready[dev] = true;
    }
    t = c.t[dev];
    return hipSuccess;
}

const uint32_t* host_mrg_table(int which_engine, int sub)
{
    const host_tables& h = host();
    if(which_engine == 32)
        return sub ? h.mrg32_sub.data() : h.mrg32_step.data();
    return sub ? h.mrg31_sub.data() : h.mrg31_step.data();
}

const uint32_t* host_mrg_radix(int which_engine)
{
    const host_tables& h = host();
    return which_engine == 32 ? h.mrg32_subr.data() : h.mrg31_subr.data();
}

const uint32_t* host_gf2_table(int xorwow, int seq)
{
    const host_tables& h = host();
    if(xorwow)
        return seq ? h.xorwow_seq.data() : h.xorwow_jump.data();
    return seq ? h.lfsr_seq.data() : h.lfsr_jump.data();
}

} // namespace rr
