// Seeded synthetic op-stream generators (device side), include/lslab_workload.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lslab_c.h"
#include "../../include/lslab_workload.h"

namespace {

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// splitmix64 step as a counter-based generator: uniform 64-bit word for
// (seed, stream index, draw).
__device__ __forceinline__ uint64_t rnd(uint64_t seed, uint64_t i, uint64_t draw) {
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i * 8ull + draw + 1ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

// Bijection of [0, 2^b): odd multiplies and xor-shifts, each invertible mod 2^b.
__device__ __forceinline__ uint64_t perm_bits(uint64_t x, unsigned b, uint64_t seed) {
    const uint64_t mask = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
    const unsigned sh = (b + 1) / 2;
    x = (x + seed) & mask;
    x = (x * 0x9e3779b97f4a7c15ull) & mask;
    x ^= x >> sh;
    x = (x * 0xbf58476d1ce4e5b9ull) & mask;
    x ^= x >> sh;
    x = (x * 0x94d049bb133111ebull) & mask;
    x ^= x >> sh;
    return x;
}

// Rejection-inversion sampling of Zipf(s, N) (W. Hoermann, G. Derflinger,
// "Rejection-inversion to generate variates from monotone discrete
// distributions", ACM TOMACS 6(3), 1996), with the hat H(x) = integral of
// x^-s written through expm1/log1p so that s -> 1 stays accurate.
struct Zipf {
    double s, hx1, hN, sv;
    uint64_t N;
    __device__ static double helper1(double y) {  // log1p(y)/y
        return fabs(y) > 1e-8 ? log1p(y) / y : 1.0 - y * (0.5 - y * (1.0 / 3.0 - 0.25 * y));
    }
    __device__ static double helper2(double y) {  // expm1(y)/y
        return fabs(y) > 1e-8 ? expm1(y) / y : 1.0 + y * 0.5 * (1.0 + y * (1.0 / 3.0) * (1.0 + 0.25 * y));
    }
    __device__ double H(double x) const {
        const double lx = log(x);
        return helper2((1.0 - s) * lx) * lx;
    }
    __device__ double Hinv(double u) const {
        double t = u * (1.0 - s);
        if (t < -1.0) t = -1.0;
        return exp(helper1(t) * u);
    }
    __device__ double h(double x) const { return exp(-s * log(x)); }
    __device__ void init(double s_, uint64_t N_) {
        s = s_;
        N = N_;
        hx1 = H(1.5) - 1.0;
        hN = H((double)N + 0.5);
        sv = 2.0 - Hinv(H(2.5) - h(2.0));
    }
    __device__ uint64_t sample(uint64_t seed, uint64_t i) const {
        for (uint64_t d = 0; d < 64; ++d) {
            const double u = hN + u01(rnd(seed, i, d)) * (hx1 - hN);
            const double x = Hinv(u);
            double kd = floor(x + 0.5);
            if (kd < 1.0) kd = 1.0;
            if (kd > (double)N) kd = (double)N;
            const uint64_t k = (uint64_t)kd;
            if (kd - x <= sv || u >= H(kd + 0.5) - h(kd)) return k;
        }
        return 1;  // acceptance probability per draw is > 0.5; 64 rejections do not happen
    }
};

__global__ void k_gen_keys(char* out, int kb, uint64_t first, uint64_t n, uint64_t seed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (kb == 4) reinterpret_cast<uint32_t*>(out)[i] = fmix32((uint32_t)(first + i) ^ (uint32_t)seed);
    else reinterpret_cast<uint64_t*>(out)[i] = fmix64((first + i) ^ seed);
}

__global__ void k_gen_values(char* out, int vb, uint64_t first, uint64_t n, uint64_t seed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t v = fmix64((first + i) ^ seed);
    if (vb == 4) reinterpret_cast<uint32_t*>(out)[i] = (uint32_t)v;
    else reinterpret_cast<uint64_t*>(out)[i] = v;
}

__global__ void k_gen_ops(lslab_gen_params p, uint8_t* ops, char* keys, char* vals, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t g = p.first_op + i;
    const uint64_t N = 1ull << p.space_log2;
    uint64_t rank;
    if (p.theta > 0.0) {
        Zipf z;
        z.init(p.theta, N);
        rank = z.sample(p.seed, g);
    } else {
        rank = 1 + (rnd(p.seed, g, 0) & (N - 1));
    }
    const uint64_t key = perm_bits(rank - 1, p.space_log2, p.seed);
    const double c = u01(rnd(p.seed ^ 0x5bd1e995ull, g, 7));
    const uint8_t op = c < p.p_get ? LSLAB_OP_GET : (c < p.p_get + p.p_put ? LSLAB_OP_PUT : LSLAB_OP_REMOVE);
    ops[i] = op;
    if (p.key_bytes == 4) reinterpret_cast<uint32_t*>(keys)[i] = (uint32_t)key;
    else reinterpret_cast<uint64_t*>(keys)[i] = key;
    if (vals) {
        const uint64_t v = fmix64(g ^ p.seed);
        if (p.val_bytes == 4) reinterpret_cast<uint32_t*>(vals)[i] = (uint32_t)v;
        else reinterpret_cast<uint64_t*>(vals)[i] = v;
    }
}

__global__ void k_gen_zipf(uint64_t* out, uint64_t n, double theta, uint64_t N, uint64_t seed,
                           uint64_t first) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Zipf z;
    z.init(theta, N);
    out[i] = z.sample(seed, first + i);
}

unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int lslab_gen_keys(void* out, int key_bytes, uint64_t first, uint64_t n, uint64_t seed,
                              hipStream_t stream) {
    if ((key_bytes != 4 && key_bytes != 8) || (n && !out) || n >= (1ull << 40)) return LSLAB_ERR_INVALID;
    if (!n) return LSLAB_OK;
    hipLaunchKernelGGL(k_gen_keys, dim3(blocks_for(n)), dim3(256), 0, stream, (char*)out, key_bytes,
                       first, n, seed);
    return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
}

extern "C" int lslab_gen_values(void* out, int val_bytes, uint64_t first, uint64_t n, uint64_t seed,
                                hipStream_t stream) {
    if ((val_bytes != 4 && val_bytes != 8) || (n && !out) || n >= (1ull << 40)) return LSLAB_ERR_INVALID;
    if (!n) return LSLAB_OK;
    hipLaunchKernelGGL(k_gen_values, dim3(blocks_for(n)), dim3(256), 0, stream, (char*)out,
                       val_bytes, first, n, seed);
    return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
}

extern "C" int lslab_gen_ops(const lslab_gen_params* p, uint8_t* ops, void* keys, void* vals,
                             uint64_t n, hipStream_t stream) {
    if (!p || (n && (!ops || !keys)) || n >= (1ull << 40)) return LSLAB_ERR_INVALID;
    if (p->space_log2 < 1 || p->space_log2 > 63) return LSLAB_ERR_INVALID;
    if ((p->key_bytes != 4 && p->key_bytes != 8) || (vals && p->val_bytes != 4 && p->val_bytes != 8))
        return LSLAB_ERR_INVALID;
    if (p->key_bytes == 4 && p->space_log2 > 32) return LSLAB_ERR_INVALID;
    if (p->theta < 0.0 || p->p_get < 0.0 || p->p_put < 0.0 || p->p_get + p->p_put > 1.0 + 1e-12)
        return LSLAB_ERR_INVALID;
    if (!n) return LSLAB_OK;
    hipLaunchKernelGGL(k_gen_ops, dim3(blocks_for(n)), dim3(256), 0, stream, *p, ops, (char*)keys,
                       (char*)vals, n);
    return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
}

extern "C" int lslab_gen_zipf_ranks(uint64_t* out, uint64_t n, double theta, uint64_t space,
                                    uint64_t seed, uint64_t first, hipStream_t stream) {
    if ((n && !out) || theta <= 0.0 || space < 1 || n >= (1ull << 40)) return LSLAB_ERR_INVALID;
    if (!n) return LSLAB_OK;
    hipLaunchKernelGGL(k_gen_zipf, dim3(blocks_for(n)), dim3(256), 0, stream, out, n, theta, space,
                       seed, first);
    return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
}
