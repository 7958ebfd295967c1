// CPU oracle for the kNN photon gather.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file; only the tests, the smoke check and the CPU-baseline leg of
// the benchmark do, and only to check results / as the timed CPU reference.
//
// What it restates (paths relative to the reference checkout):
//   * distance2                       include/photon_map.h:33-41
//   * KdTree::buildNode / buildTree   include/photon_map.h:61-105, 157-164
//   * KdTree::searchKNearestNode      include/photon_map.h:107-147
//   * KdTree::searchKNearest          include/photon_map.h:166-183
//   * Lambert::evaluate               include/material.h:80-88 (via
//     Primitive::evaluateBxDF, include/primitive.h:32-42, and the local frame
//     of include/core.h:206-209,234-237: only the y = shading-normal component
//     of the local direction is ever read)
//   * computeRadianceWithPhotonMap    include/integrator.h:133-152
// and adds what has no reference counterpart but is needed to check results:
//   * brute-force exact kNN under the (FP32 d^2, photon index) order,
//   * an FP64 radiance estimate and its FP64 adjoint (gradients with respect
//     to per-photon flux and per-material albedo).
//
// Build flags matter: the file is compiled without -march (so no FMA) and with
// -ffp-contract=off, so every FP32 operation rounds exactly as the
// reference's x86-64 SSE2 build does.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.hpp"

using namespace orc;

extern "C" {

int orc_version(void) { return 1; }

// ---- kd-tree restatement ---------------------------------------------------
void* orc_kdtree_new(int64_t n, const float* px, const float* py,
                     const float* pz) {
  auto* t = new KdTreeOracle();
  t->n = static_cast<int>(n);
  t->x.assign(px, px + n);
  t->y.assign(py, py + n);
  t->z.assign(pz, pz + n);
  std::vector<int> indices(n);
  std::iota(indices.begin(), indices.end(), 0);
  t->nodes.reserve(n);
  t->build(indices.data(), t->n, 0);
  return t;
}

void orc_kdtree_free(void* h) { delete static_cast<KdTreeOracle*>(h); }

int64_t orc_kdtree_num_nodes(void* h) {
  return static_cast<int64_t>(static_cast<KdTreeOracle*>(h)->nodes.size());
}

// out_idx is m x k (rows padded with -1), out_r2 is m.
int orc_kdtree_knn(void* h, int64_t m, const float* qx, const float* qy,
                   const float* qz, int k, int32_t* out_idx, float* out_r2,
                   int nthreads) {
  if (k < 1) return 1;
  const auto* t = static_cast<KdTreeOracle*>(h);
  const int nt = resolve_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
  for (int64_t i = 0; i < m; ++i) {
    t->query(qx[i], qy[i], qz[i], k, out_idx + i * k, out_r2 + i);
  }
  return 0;
}

// ---- brute force -------------------------------------------------------------
// Exact k smallest (FP32 d^2, index) keys, written largest first like the
// reference's drain order.
int orc_brute_knn(int64_t n, const float* px, const float* py, const float* pz,
                  int64_t m, const float* qx, const float* qy, const float* qz,
                  int k, int32_t* out_idx, float* out_r2, int nthreads) {
  if (k < 1) return 1;
  const int nt = resolve_threads(nthreads);
#pragma omp parallel num_threads(nt)
  {
    std::vector<std::pair<float, int>> keys(n);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < m; ++i) {
      for (int64_t j = 0; j < n; ++j) {
        keys[j] = {dist2_f32(qx[i], qy[i], qz[i], px[j], py[j], pz[j]),
                   static_cast<int>(j)};
      }
      const int64_t cnt = std::min<int64_t>(k, n);
      std::partial_sort(keys.begin(), keys.begin() + cnt, keys.end());
      int32_t* row = out_idx + i * k;
      for (int64_t j = 0; j < cnt; ++j) row[j] = keys[cnt - 1 - j].second;
      for (int64_t j = cnt; j < k; ++j) row[j] = -1;
      out_r2[i] = cnt > 0 ? keys[cnt - 1].first : 0.0f;
    }
  }
  return 0;
}

// ---- radiance estimates ------------------------------------------------------
// FP64 estimate on a given neighbour set.  The discrete inputs (set
// membership, r^2 and the two FP32 sign tests) are taken exactly as the
// reference computes them; everything else is carried in double.
//   L_c = sum_i s_i * (rho_c / pi) * flux_{i,c} / (n_emit * pi * r^2)
// with pi = the reference's float PI.
int orc_radiance_fp64(int64_t n, const float* wi_soa, const float* flux_soa,
                      int64_t m, const void* queries, const void* materials,
                      int n_mats, int k, const int32_t* idx, const float* r2,
                      double n_emit, double* out_rgb, int nthreads) {
  const auto* q = static_cast<const QueryRec*>(queries);
  const auto* mats = static_cast<const MaterialRec*>(materials);
  const double pi = static_cast<double>(kPiF);
  const int nt = resolve_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < m; ++i) {
    const QueryRec& qr = q[i];
    double sum[3] = {0.0, 0.0, 0.0};
    int count = 0;
    const bool lambert = qr.material >= 0 && qr.material < n_mats &&
                         mats[qr.material].type == 0;
    const float cos_o = dot3_f32(qr.wo, qr.ns);
    for (int j = 0; j < k; ++j) {
      const int64_t p = idx[i * k + j];
      if (p < 0) continue;
      ++count;
      const float wi[3] = {wi_soa[p], wi_soa[n + p], wi_soa[2 * n + p]};
      if (!(lambert && lambert_passes(cos_o, dot3_f32(wi, qr.ns)))) continue;
      for (int c = 0; c < 3; ++c) sum[c] += static_cast<double>(flux_soa[c * n + p]);
    }
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      if (count > 0 && lambert) {
        v = (static_cast<double>(mats[qr.material].rho[c]) / pi) * sum[c] /
            (n_emit * pi * static_cast<double>(r2[i]));
      }
      out_rgb[i * 3 + c] = v;
    }
  }
  return 0;
}

// FP32 estimate exactly as the reference evaluates it, on a given
// (farthest-first) neighbour list.
int orc_radiance_ref32(int64_t n, const float* wi_soa, const float* flux_soa,
                       int64_t m, const void* queries, const void* materials,
                       int n_mats, int k, const int32_t* idx, const float* r2,
                       double n_emit, float* out_rgb, int nthreads) {
  const auto* q = static_cast<const QueryRec*>(queries);
  const auto* mats = static_cast<const MaterialRec*>(materials);
  const int nt = resolve_threads(nthreads);
  const float n_emit_f = static_cast<float>(n_emit);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < m; ++i) {
    int count = 0;
    while (count < k && idx[i * k + count] >= 0) ++count;
    estimate_ref32(wi_soa, flux_soa, n, q[i], mats, n_mats, idx + i * k, count,
                   r2[i], n_emit_f, out_rgb + i * 3);
  }
  return 0;
}

// The whole reference gather on the CPU: kd-tree search then the FP32
// estimate, per query, OpenMP dynamic schedule over queries.  This is the
// CPU baseline the benchmark times.
int orc_gather_ref32(void* h, const float* wi_soa, const float* flux_soa,
                     int64_t m, const void* queries, const void* materials,
                     int n_mats, int k, double n_emit, float* out_rgb,
                     float* out_r2, int nthreads) {
  if (k < 1) return 1;
  const auto* t = static_cast<KdTreeOracle*>(h);
  const auto* q = static_cast<const QueryRec*>(queries);
  const auto* mats = static_cast<const MaterialRec*>(materials);
  const int nt = resolve_threads(nthreads);
  const float n_emit_f = static_cast<float>(n_emit);
#pragma omp parallel num_threads(nt)
  {
    std::vector<int32_t> row(k);
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < m; ++i) {
      float r2 = 0.0f;
      const int cnt =
          t->query(q[i].pos[0], q[i].pos[1], q[i].pos[2], k, row.data(), &r2);
      estimate_ref32(wi_soa, flux_soa, t->n, q[i], mats, n_mats, row.data(),
                     cnt, r2, n_emit_f, out_rgb + i * 3);
      out_r2[i] = r2;
    }
  }
  return 0;
}

// FP64 adjoint of orc_radiance_fp64 on a given neighbour set.
//   dL/dflux_{p,c}  += g_{q,c} * s_qp * (rho_c / pi) / (n_emit * pi * r_q^2)
//   dL/drho_{m,c}   += g_{q,c} * sum_p s_qp flux_{p,c} / (pi * n_emit * pi * r_q^2)
// Outputs are accumulated (caller zeroes them).  dflux is 3 x n (SoA), drho
// is n_mats x 3.  Serial: the scatter order is then fixed.
int orc_adjoint_fp64(int64_t n, const float* wi_soa, const float* flux_soa,
                     int64_t m, const void* queries, const void* materials,
                     int n_mats, int k, const int32_t* idx, const float* r2,
                     double n_emit, const float* dl_drgb, double* dflux,
                     double* drho) {
  const auto* q = static_cast<const QueryRec*>(queries);
  const auto* mats = static_cast<const MaterialRec*>(materials);
  const double pi = static_cast<double>(kPiF);
  for (int64_t i = 0; i < m; ++i) {
    const QueryRec& qr = q[i];
    const bool lambert = qr.material >= 0 && qr.material < n_mats &&
                         mats[qr.material].type == 0;
    if (!lambert) continue;
    int count = 0;
    for (int j = 0; j < k; ++j) count += idx[i * k + j] >= 0;
    if (count == 0) continue;
    const float cos_o = dot3_f32(qr.wo, qr.ns);
    const double denom = n_emit * pi * static_cast<double>(r2[i]);
    double sum[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < k; ++j) {
      const int64_t p = idx[i * k + j];
      if (p < 0) continue;
      const float wi[3] = {wi_soa[p], wi_soa[n + p], wi_soa[2 * n + p]};
      if (!lambert_passes(cos_o, dot3_f32(wi, qr.ns))) continue;
      for (int c = 0; c < 3; ++c) {
        const double g = static_cast<double>(dl_drgb[i * 3 + c]);
        dflux[c * n + p] +=
            g * (static_cast<double>(mats[qr.material].rho[c]) / pi) / denom;
        sum[c] += static_cast<double>(flux_soa[c * n + p]);
      }
    }
    for (int c = 0; c < 3; ++c) {
      const double g = static_cast<double>(dl_drgb[i * 3 + c]);
      drho[qr.material * 3 + c] += g * sum[c] / (pi * denom);
    }
  }
  return 0;
}

}  // extern "C"
