// particles.hip -- spatial ordering of the forward particle set.
//
// A forward particle (lp_motion_m.f90:15-54) is advected from its origin node
// for a whole h interval; as the set disperses, a wavefront of particles that
// started on neighbouring nodes ends up spread over the domain and its
// interpolation gathers stop sharing cache lines.  Every particle's update
// depends on its own state only, so the set may be stored in any order: it is
// radix-sorted by the slab-order index of its tracked node (the order the
// semi-Lagrangian step walks the nodes in), which makes the 64 particles of a
// wavefront gather from a few neighbouring grid rows again and keeps what one
// XCD has in flight within its L2.
// The origin index (lp%no0, data_m.f90:392-396) travels with each particle,
// and the displacement is scattered back to origin order wherever the
// reference reads it by origin (exchange_lpmap, comms_m.f90:1191-1376).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "device_common.h"
#include "lcs_device.h"

namespace lcs {

constexpr int kPBlock = 256;

__global__ void __launch_bounds__(kPBlock) k_sort_keys(int n0, int n1, int n2, int np, const int16_t* __restrict__ ns0,
                                                       const int16_t* __restrict__ ns1,
                                                       const int16_t* __restrict__ ns2, unsigned* __restrict__ key,
                                                       unsigned* __restrict__ val) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  key[q] = (unsigned)encode_interior_slab(n0, n1, n2, ns0[q], ns1[q], ns2[q]);
  val[q] = (unsigned)q;
}

template <typename T>
__global__ void __launch_bounds__(kPBlock) k_permute(int np, const unsigned* __restrict__ perm, ParticleSet<T> a,
                                                     ParticleSet<T> b) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const unsigned s = perm[q];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    b.x[d][q] = a.x[d][s];
    b.d[d][q] = a.d[d][s];
    b.xl[d][q] = a.xl[d][s];
    b.dl[d][q] = a.dl[d][s];
    b.n[d][q] = a.n[d][s];
  }
  b.org[q] = a.org[s];
}

template <typename T>
__global__ void __launch_bounds__(kPBlock) k_unpermute(int np, const int* __restrict__ org, CF3<T> src, CF3<T> lo,
                                                       F3<T> dst) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int o = org[q];
#pragma unroll
  for (int d = 0; d < 3; d++) dst.c[d][o] = lo.c[d] ? src.c[d][q] + lo.c[d][q] : src.c[d][q];
}

static int key_bits(long long np) {
  int b = 1;
  while (b < 31 && (1LL << b) < np) b++;
  return b;
}

size_t particle_sort_temp_bytes(long long np, long long nkeys) {
  size_t bytes = 0;
  hipcub::DoubleBuffer<unsigned> k(nullptr, nullptr), v(nullptr, nullptr);
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, (int)np, 0, key_bits(nkeys)) != hipSuccess) return 0;
  return bytes;
}

// the keys index the flow grid g (tracked nodes); the np particles may live
// on another grid
template <typename T>
int launch_particle_sort(const GridDev<T>& g, ParticleSet<T> a, ParticleSet<T> b, long long npl, unsigned* keys,
                         unsigned* vals, void* temp, size_t temp_bytes, hipStream_t s) {
  const int np = (int)npl;
  const dim3 grid((np + kPBlock - 1) / kPBlock);
  hipLaunchKernelGGL(k_sort_keys, grid, dim3(kPBlock), 0, s, g.n[0], g.n[1], g.n[2], np, a.n[0], a.n[1], a.n[2], keys,
                     vals);
  hipcub::DoubleBuffer<unsigned> kb(keys, keys + np), vb(vals, vals + np);
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, kb, vb, np, 0, key_bits(g.np), s) != hipSuccess) return 1;
  hipLaunchKernelGGL((k_permute<T>), grid, dim3(kPBlock), 0, s, np, (const unsigned*)vb.Current(), a, b);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// brick-key sort of the re-ordered composition (kernels.hip k_compose_slot)
size_t compose_sort_temp_bytes(long long np, long long nkeys) {
  size_t bytes = 0;
  hipcub::DoubleBuffer<unsigned> k(nullptr, nullptr), v(nullptr, nullptr);
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, (int)np, 0, key_bits(nkeys)) != hipSuccess) return 0;
  return bytes;
}

const unsigned* launch_compose_sort(unsigned* keys, unsigned* vals, long long npl, long long nkeys, void* temp,
                                    size_t temp_bytes, hipStream_t s) {
  const int np = (int)npl;
  hipcub::DoubleBuffer<unsigned> kb(keys, keys + np), vb(vals, vals + np);
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, kb, vb, np, 0, key_bits(nkeys), s) != hipSuccess) return nullptr;
  return vb.Current();
}

template <typename T>
void launch_unpermute(long long npl, const int* org, CF3<T> src, const T* const lo[3], F3<T> dst, hipStream_t s) {
  const int np = (int)npl;
  CF3<T> l{{lo ? lo[0] : nullptr, lo ? lo[1] : nullptr, lo ? lo[2] : nullptr}};
  hipLaunchKernelGGL((k_unpermute<T>), dim3((np + kPBlock - 1) / kPBlock), dim3(kPBlock), 0, s, np, org, src, l,
                     dst);
}

template int launch_particle_sort<float>(const GridDev<float>&, ParticleSet<float>, ParticleSet<float>, long long,
                                         unsigned*, unsigned*, void*, size_t, hipStream_t);
template int launch_particle_sort<double>(const GridDev<double>&, ParticleSet<double>, ParticleSet<double>,
                                          long long, unsigned*, unsigned*, void*, size_t, hipStream_t);
template void launch_unpermute<float>(long long, const int*, CF3<float>, const float* const*, F3<float>,
                                      hipStream_t);
template void launch_unpermute<double>(long long, const int*, CF3<double>, const double* const*, F3<double>,
                                       hipStream_t);

}  // namespace lcs
