// Becke fuzzy-cell partition of an atom-centred integration grid.
//
// Reference computation replaced: the partition weights of PySCF's dft.Grids, which the
// reference takes as mf.grids.weights (grad_dft/interface/pyscf.py:54-60, grid_from_pyscf).
// For a point r owned by atom o (it came from o's radial x angular rule with quadrature
// volume v):
//   w(r) = v * P_o(r) / sum_i P_i(r),   P_i(r) = prod_{j != i} s(nu_ij(r)),
//   mu_ij = (|r - R_i| - |r - R_j|) / |R_i - R_j|,   nu_ij = mu_ij + a_ij (1 - mu_ij^2),
//   s(nu) = (1 - f(f(f(nu)))) / 2,   f(x) = (3 - x^2) x / 2,
// with the size-adjustment matrix a (antisymmetric, |a_ij| <= 1/2) given by the caller.
//
// One thread per point; atoms, a and the inter-atomic distances are read with wave-uniform
// indices (scalar loads).  The pair function is evaluated once per ordered pair with the
// higher atom index first (s for the lower atom is 1 - s of the higher one, exactly: nu and
// the f iterations are odd), the factors of each P_i multiplied in increasing j and the P_i
// summed in increasing i, and floating-point contraction is off, so the weights equal a
// sequential double-precision evaluation of the same formulas operation by operation.
#include <hip/hip_runtime.h>

#include "gdx_common.h"
#include "gdx_profile.h"

namespace gdx {
namespace {

constexpr int BK_TPB = 256;

__device__ __forceinline__ double point_distance(double x, double y, double z, const double* __restrict__ c) {
#pragma clang fp contract(off)
  const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

__global__ __launch_bounds__(BK_TPB) void becke_kernel(int npts, const double* __restrict__ coords,
                                                       const int* __restrict__ owner,
                                                       const double* __restrict__ vol, int natm,
                                                       const double* __restrict__ atoms,
                                                       const double* __restrict__ adist,
                                                       const double* __restrict__ adj, double* __restrict__ w) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * BK_TPB + threadIdx.x;
  if (r >= npts) return;
  const double x = coords[3 * r], y = coords[3 * r + 1], z = coords[3 * r + 2];
  const int o = owner[r];
  double total = 0.0, po = 0.0;
  for (int i = 0; i < natm; ++i) {
    const double di = point_distance(x, y, z, atoms + 3 * i);
    double p = 1.0;
    for (int j = 0; j < natm; ++j) {
      if (j == i) continue;
      const double dj = point_distance(x, y, z, atoms + 3 * j);
      const int hi = i > j ? i : j, lo = i > j ? j : i;
      const double dh = i > j ? di : dj, dl = i > j ? dj : di;
      double g = (dh - dl) / adist[hi * natm + lo];
      g = g + adj[hi * natm + lo] * (1.0 - g * g);
#pragma unroll
      for (int it = 0; it < 3; ++it) g = (3.0 - g * g) * g * 0.5;
      p *= i > j ? 0.5 * (1.0 - g) : 0.5 * (1.0 + g);
    }
    total += p;
    if (i == o) po = p;
  }
  w[r] = vol[r] * (po / total);
}

}  // namespace
}  // namespace gdx

using namespace gdx;

extern "C" int gdx_becke_weights(void* stream, int32_t npts, const double* coords, const int32_t* owner,
                                 const double* vol, int32_t natm, const double* atoms, const double* adist,
                                 const double* adjust, double* weights) {
  GDX_REQUIRE(npts >= 0 && natm > 0, "gdx_becke_weights: bad sizes");
  GDX_REQUIRE(npts == 0 || (coords && owner && vol && atoms && weights), "gdx_becke_weights: null pointer");
  GDX_REQUIRE(natm == 1 || (adist && adjust), "gdx_becke_weights: null atom-pair matrices");
  if (npts == 0) return GDX_OK;
  GDX_LAUNCH("becke_weights", becke_kernel, dim3((unsigned)((npts + BK_TPB - 1) / BK_TPB)), dim3(BK_TPB), 0,
             (hipStream_t)stream, npts, coords, (const int*)owner, vol, natm, atoms, adist, adjust, weights);
  GDX_HIP_CHECK(hipGetLastError());
  return GDX_OK;
}
