// Sort-by-tet and compaction of the particle arrays.
//
// Keys: owning tet for live particles, n_tets for live particles outside the
// mesh, n_tets + 1 for settled ones.  A radix sort of (key, index) followed by
// one gather of every particle array puts particles of a tet next to each
// other (coalesced tet/E/map loads, few distinct nodes per block for the
// charge reduction) and moves settled particles behind the live ones; the
// step kernels then only cover the live prefix (n_active).
#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace sg {

__global__ void make_keys_kernel(const int32_t* __restrict__ tet, const int32_t* __restrict__ flags, int64_t n,
                                 int32_t n_tets, int32_t* __restrict__ keys, int32_t* __restrict__ idx)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    int32_t f = flags[j], t = tet[j];
    keys[j] = (f & kSettled) ? n_tets + 1 : (t < 0 ? n_tets : t);
    idx[j] = (int32_t)j;
}

template <typename R>
__global__ void permute_kernel(DevParticles<R> src, R* x, R* y, R* z, double* vx, double* vy, double* vz, int32_t* tet,
                               int32_t* flags, int32_t* hit, uint32_t* pidx, const int32_t* __restrict__ perm,
                               const int32_t* __restrict__ keys_sorted, int32_t n_tets, int64_t n,
                               int64_t* __restrict__ n_active)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    int64_t s = perm[j];
    x[j] = src.x[s];
    y[j] = src.y[s];
    z[j] = src.z[s];
    vx[j] = src.vx[s];
    vy[j] = src.vy[s];
    vz[j] = src.vz[s];
    tet[j] = src.tet[s];
    flags[j] = src.flags[s];
    hit[j] = src.hit[s];
    pidx[j] = src.pidx[s];
    // the live prefix ends where the settled keys begin
    bool live = keys_sorted[j] <= n_tets;
    bool next_live = (j + 1 < n) && keys_sorted[j + 1] <= n_tets;
    if (j == 0 && !live) *n_active = 0;
    if (live && !next_live) *n_active = j + 1;
}

size_t sort_temp_bytes(int64_t n, int nbits)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                             (const int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0, nbits);
    return bytes;
}

template <typename R>
hipError_t sort_particles(hipStream_t s, const DevParticles<R>& P, R* x, R* y, R* z, double* vx, double* vy, double* vz,
                          int32_t* tet, int32_t* flags, int32_t* hit, uint32_t* pidx, int32_t* keys,
                          int32_t* keys_out, int32_t* idx, int32_t* perm, void* temp, size_t temp_bytes,
                          int32_t n_tets, int nbits, int64_t* n_active)
{
    int64_t n = P.n;
    if (n == 0) return hipSuccess;
    unsigned nb = (unsigned)((n + 255) / 256);
    make_keys_kernel<<<nb, 256, 0, s>>>(P.tet, P.flags, n, n_tets, keys, idx);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_out, idx, perm, (int)n, 0, nbits, s);
    if (e != hipSuccess) return e;
    permute_kernel<R><<<nb, 256, 0, s>>>(P, x, y, z, vx, vy, vz, tet, flags, hit, pidx, perm, keys_out, n_tets, n,
                                         n_active);
    return hipGetLastError();
}

template hipError_t sort_particles<double>(hipStream_t, const DevParticles<double>&, double*, double*, double*,
                                           double*, double*, double*, int32_t*, int32_t*, int32_t*, uint32_t*,
                                           int32_t*, int32_t*, int32_t*, int32_t*, void*, size_t, int32_t, int,
                                           int64_t*);
template hipError_t sort_particles<float>(hipStream_t, const DevParticles<float>&, float*, float*, float*, double*,
                                          double*, double*, int32_t*, int32_t*, int32_t*, uint32_t*, int32_t*,
                                          int32_t*, int32_t*, int32_t*, void*, size_t, int32_t, int, int64_t*);

}  // namespace sg
