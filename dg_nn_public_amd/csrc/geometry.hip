// Affine element and face geometry on the device (gfx950): the per-element quantities the
// operator kernels read, formed from the P3 node coordinates in one pass.  Restates the
// reference's Jacobians, determinants, inverse and face normals
// (z02/shape_function.py:1321-1384 volume, :1730-1934 faces: normal and area from the
// cross product of J . drst_duv, h = sqrt(sum sdetwei) in 3D) for straight-sided
// elements, where J is constant: J[i][j] = x_i[j] - x_D[j] (rows are edge vectors from
// vertex D), and checks that every P3 node sits at its affine position.
#include <algorithm>
#include <climits>
#include <cmath>

#include "dgmi_internal.h"

namespace dgmi {

// element tables the kernel needs, passed by value
struct GeoTabs {
  double P[20 * 4];      // P1 interpolation at the P3 nodes: x_n = sum_k P[n][k] v_k
  double drst[4 * 3 * 2];  // d(xi)/d(face params) per face (3D: [f][i][k]; 2D: [f][i])
  double swsum;          // sum of the face quadrature weights
  double eta;            // penalty coefficient
  int face_kappa;        // 0 arithmetic mean, 1 harmonic mean on interior faces
  int opp[4];            // the vertex opposite face f
};

constexpr int GEO_THREADS = 256;

// One thread per element.  Writes (structure of arrays over elements):
//   G[c*nele + e]          kappa detJ (invJ^T invJ) components (3D: 00 11 22 01 02 12)
//   cf[(f*D + j)*nele + e] invJ^T n_f
//   wf[f*nele + e]         kappa_f |face| (kappa_f = 1 on Neumann faces, 0 without a term)
//   pen[f*nele + e]        eta / h_f
//   nbface, fcode          neighbour face id / code | alignment << 4 | frame << 8
//   cnb[(f*D + j)*nele + e] invJ'^T n' of the neighbour across face f in this face's frame
//   detJ[e]
// The frame of an interior face f: the neighbour's vertices relabelled so that the shared
// vertices keep this element's local numbers and its fourth (third) vertex takes the
// number of this element's vertex opposite f.  sigma[k] = the neighbour's own number of
// relabelled vertex k, matched by position; frame = the lexicographic rank of sigma among
// the permutations of D+1 vertices (0 = identity); J' = its Jacobian in that numbering,
// n' = -n_f its outward normal.  *badface = the smallest face whose vertices do not match.
// blockmax[block] = max over the block's elements of the node deviation from the affine
// position relative to the element size; *bad = the smallest degenerate element index.
template <int D>
__global__ __launch_bounds__(GEO_THREADS) void k_geometry(
    int nele, const double* __restrict__ x, const double* __restrict__ kappa,
    const int32_t* __restrict__ nbf, const int32_t* __restrict__ bct,
    const int32_t* __restrict__ alnmt, GeoTabs t, double* __restrict__ G,
    double* __restrict__ cf, double* __restrict__ wf, double* __restrict__ pen,
    int32_t* __restrict__ nbface, int32_t* __restrict__ fcode, double* __restrict__ detJ,
    double* __restrict__ cnb, double* __restrict__ blockmax, int* __restrict__ bad,
    int* __restrict__ badface) {
  constexpr int NL = D == 3 ? 20 : 10, NB = D + 1, NF = D + 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double dev = 0.0;
  if (e < nele) {
    const double* xe = x + (size_t)e * D * NL;  // (D, NL)
    double v[4][3] = {};
    for (int k = 0; k < NB; ++k)
      for (int j = 0; j < D; ++j) v[k][j] = xe[j * NL + k];
    double J[3][3] = {};
    double hs = 0.0;
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        J[i][j] = v[i][j] - v[D][j];
        hs = fmax(hs, fabs(J[i][j]));
      }
    for (int n = 0; n < NL; ++n)
      for (int j = 0; j < D; ++j) {
        double ex = 0.0;
        for (int k = 0; k < NB; ++k) ex += t.P[n * 4 + k] * v[k][j];
        dev = fmax(dev, fabs(xe[j * NL + n] - ex) / (hs > 0.0 ? hs : 1.0));
      }
    double det, inv[3][3] = {};
    if constexpr (D == 3) {
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
            J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    } else {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      inv[0][0] = J[1][1] / det;
      inv[0][1] = -J[0][1] / det;
      inv[1][0] = -J[1][0] / det;
      inv[1][1] = J[0][0] / det;
    }
    if (!(fabs(det) > 0.0) || !isfinite(det)) {
      atomicMin(bad, e);
      det = 1.0;  // keep the stores below finite; the context is refused
    }
    detJ[e] = det;
    const double ke = kappa[e];
    auto Gjk = [&](int j, int k) {
      double s = 0.0;
      for (int i = 0; i < D; ++i) s += inv[i][j] * inv[i][k];
      return ke * det * s;
    };
    if constexpr (D == 3) {
      G[0 * (size_t)nele + e] = Gjk(0, 0);
      G[1 * (size_t)nele + e] = Gjk(1, 1);
      G[2 * (size_t)nele + e] = Gjk(2, 2);
      G[3 * (size_t)nele + e] = Gjk(0, 1);
      G[4 * (size_t)nele + e] = Gjk(0, 2);
      G[5 * (size_t)nele + e] = Gjk(1, 2);
    } else {
      G[0 * (size_t)nele + e] = Gjk(0, 0);
      G[1 * (size_t)nele + e] = Gjk(1, 1);
      G[2 * (size_t)nele + e] = Gjk(0, 1);
    }
    for (int f = 0; f < NF; ++f) {
      double nrm[3] = {0.0, 0.0, 0.0}, sdet;
      if constexpr (D == 3) {
        double tg[2][3];
        for (int k = 0; k < 2; ++k)
          for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i) s += J[i][j] * t.drst[(f * 3 + i) * 2 + k];
            tg[k][j] = s;
          }
        const double cr[3] = {tg[0][1] * tg[1][2] - tg[0][2] * tg[1][1],
                              tg[0][2] * tg[1][0] - tg[0][0] * tg[1][2],
                              tg[0][0] * tg[1][1] - tg[0][1] * tg[1][0]};
        sdet = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int j = 0; j < 3; ++j) nrm[j] = cr[j] / sdet;
      } else {
        double tg[2];
        for (int j = 0; j < 2; ++j) {
          double s = 0.0;
          for (int i = 0; i < 2; ++i) s += J[i][j] * t.drst[f * 2 + i];
          tg[j] = s;
        }
        sdet = sqrt(tg[0] * tg[0] + tg[1] * tg[1]);
        nrm[0] = tg[1] / sdet;
        nrm[1] = -tg[0] / sdet;
      }
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
        for (int i = 0; i < D; ++i) s += inv[i][j] * nrm[i];
        cf[(size_t)(f * D + j) * nele + e] = s;
      }
      double h = sdet * t.swsum;
      if constexpr (D == 3) h = sqrt(h);
      pen[(size_t)f * nele + e] = t.eta / h;
      const int z = e * NF + f;
      const int type = bct[z];
      const int nb = nbf[z];
      double kf = 0.0;
      int code = 2, nbz = -1;
      double cn[3] = {0.0, 0.0, 0.0};
      if (type < 0 && nb >= 0) {
        const double kn = kappa[nb / NF];
        kf = t.face_kappa == 1 ? 2.0 * ke * kn / (ke + kn) : 0.5 * (ke + kn);
        nbz = nb;
        // the neighbour in this face's frame
        const double* xn = x + (size_t)(nb / NF) * D * NL;
        double vn[4][3] = {};
        for (int k = 0; k < NB; ++k)
          for (int j = 0; j < D; ++j) vn[k][j] = xn[j * NL + k];
        int sig[4] = {0, 0, 0, 0}, used = 0;
        bool ok = true;
        for (int k = 0; k < NB; ++k) {
          if (k == t.opp[f]) continue;
          int best = 0;
          double bd = INFINITY;
          for (int q = 0; q < NB; ++q) {
            double d2 = 0.0;
            for (int j = 0; j < D; ++j) d2 += (vn[q][j] - v[k][j]) * (vn[q][j] - v[k][j]);
            if (d2 < bd) {
              bd = d2;
              best = q;
            }
          }
          ok = ok && !((used >> best) & 1);
          used |= 1 << best;
          sig[k] = best;
        }
        for (int q = 0; q < NB; ++q)
          if (!((used >> q) & 1)) sig[t.opp[f]] = q;
        int rank = 0;
        for (int i = 0; i < NB; ++i) {
          int less = 0;
          for (int j = i + 1; j < NB; ++j) less += sig[j] < sig[i];
          rank = rank * (NB - i) + less;
        }
        if (!ok) {
          atomicMin(badface, z);
          rank = 0;
        }
        double Jn[3][3] = {}, in[3][3] = {};
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j) Jn[i][j] = vn[sig[i]][j] - vn[sig[D]][j];
        if constexpr (D == 3) {
          const double dn = Jn[0][0] * (Jn[1][1] * Jn[2][2] - Jn[1][2] * Jn[2][1]) -
                            Jn[0][1] * (Jn[1][0] * Jn[2][2] - Jn[1][2] * Jn[2][0]) +
                            Jn[0][2] * (Jn[1][0] * Jn[2][1] - Jn[1][1] * Jn[2][0]);
          in[0][0] = (Jn[1][1] * Jn[2][2] - Jn[1][2] * Jn[2][1]) / dn;
          in[0][1] = (Jn[0][2] * Jn[2][1] - Jn[0][1] * Jn[2][2]) / dn;
          in[0][2] = (Jn[0][1] * Jn[1][2] - Jn[0][2] * Jn[1][1]) / dn;
          in[1][0] = (Jn[1][2] * Jn[2][0] - Jn[1][0] * Jn[2][2]) / dn;
          in[1][1] = (Jn[0][0] * Jn[2][2] - Jn[0][2] * Jn[2][0]) / dn;
          in[1][2] = (Jn[0][2] * Jn[1][0] - Jn[0][0] * Jn[1][2]) / dn;
          in[2][0] = (Jn[1][0] * Jn[2][1] - Jn[1][1] * Jn[2][0]) / dn;
          in[2][1] = (Jn[0][1] * Jn[2][0] - Jn[0][0] * Jn[2][1]) / dn;
          in[2][2] = (Jn[0][0] * Jn[1][1] - Jn[0][1] * Jn[1][0]) / dn;
        } else {
          const double dn = Jn[0][0] * Jn[1][1] - Jn[0][1] * Jn[1][0];
          in[0][0] = Jn[1][1] / dn;
          in[0][1] = -Jn[0][1] / dn;
          in[1][0] = -Jn[1][0] / dn;
          in[1][1] = Jn[0][0] / dn;
        }
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
          for (int i = 0; i < D; ++i) s -= in[i][j] * nrm[i];
          cn[j] = isfinite(s) ? s : 0.0;
        }
        code = (int)(alnmt[z] << 4) | (rank << 8);
      } else if (type == 0) {
        kf = ke;
        code = 1;
      } else if (type == 1) {
        // Neumann: no operator term; the right-hand side adds int v g_N without kappa
        // (z02/volume_mf_diff.py:360-400)
        kf = 1.0;
        code = 3;
      }
      nbface[z] = nbz;
      fcode[z] = code;
      for (int j = 0; j < D; ++j) cnb[(size_t)(f * D + j) * nele + e] = cn[j];
      wf[(size_t)f * nele + e] = kf * sdet;
    }
  }
  // block maximum of the deviation
  __shared__ double sh[GEO_THREADS];
  sh[threadIdx.x] = dev;
  __syncthreads();
  for (int w = GEO_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blockmax[blockIdx.x] = sh[0];
}

template <typename T>
static int to_dev(dgmi_ctx* c, const std::vector<T>& h, T** d, hipStream_t s) {
  *d = static_cast<T*>(dev_alloc(sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!*d) return -3;
  return h2d(c, *d, h.data(), sizeof(T) * h.size(), s);
}

int build_geometry_device(dgmi_ctx* c, const double* const* x_parts, const int64_t* part_off,
                          int nparts, int face_kappa, double* max_dev, int64_t* bad_elem) {
  const int D = c->D, NF = c->nface;
  const int64_t nele = c->nele, nz = nele * NF;
  hipStream_t s = c->ustream;
  DevGeom<double>& g = c->g64;
  const size_t ng = (size_t)c->ng * nele;
  g.G = static_cast<double*>(c->alloc(sizeof(double) * ng));
  g.cf = static_cast<double*>(c->alloc(sizeof(double) * nz * D));
  g.wf = static_cast<double*>(c->alloc(sizeof(double) * nz));
  g.pen = static_cast<double*>(c->alloc(sizeof(double) * nz));
  c->nbface = static_cast<int32_t*>(c->alloc(sizeof(int32_t) * nz));
  c->fcode = static_cast<int32_t*>(c->alloc(sizeof(int32_t) * nz));
  c->detJ = static_cast<double*>(c->alloc(sizeof(double) * nele));
  g.cnb = static_cast<double*>(c->alloc(sizeof(double) * nz * D));
  if (!g.G || !g.cf || !g.wf || !g.pen || !c->nbface || !c->fcode || !c->detJ || !g.cnb) {
    set_error(alloc_failure("geometry"));
    return -3;
  }
  GeoTabs t{};
  for (int n = 0; n < c->nloc; ++n)
    for (int k = 0; k <= D; ++k) t.P[n * 4 + k] = c->el.p1_prolong[n * (D + 1) + k];
  for (size_t i = 0; i < c->el.drst.size() && i < 24; ++i) t.drst[i] = c->el.drst[i];
  t.swsum = 0.0;
  for (double w : c->el.sweight) t.swsum += w;
  t.eta = c->eta_e;
  t.face_kappa = face_kappa;
  for (int f = 0; f < NF; ++f) {
    int mask = 0;
    for (int i = 0; i < D; ++i) mask |= 1 << c->el.face_verts[f * D + i];
    for (int k = 0; k <= D; ++k)
      if (!((mask >> k) & 1)) t.opp[f] = k;
  }
  // node relabellings of the face frames: for the vertex permutation sigma of rank r
  // (lexicographic), node n of the relabelled element is the node m of the original
  // whose barycentric weights satisfy w_m[sigma[k]] = w_n[k] for every vertex k
  {
    const int NB = D + 1, NL = c->nloc;
    std::vector<int> sig(NB);
    for (int k = 0; k < NB; ++k) sig[k] = k;
    std::vector<uint8_t> perm;
    const double* w = c->el.p1_prolong.data();  // (nloc, D+1)
    do {
      for (int n = 0; n < NL; ++n) {
        int found = -1;
        for (int m = 0; m < NL && found < 0; ++m) {
          bool eq = true;
          for (int k = 0; k < NB; ++k) eq = eq && std::fabs(w[m * NB + sig[k]] - w[n * NB + k]) < 1e-12;
          if (eq) found = m;
        }
        if (found < 0) {
          set_error("P3 nodes are not closed under vertex permutations");
          return -2;
        }
        perm.push_back((uint8_t)found);
      }
    } while (std::next_permutation(sig.begin(), sig.end()));
    c->nperm = static_cast<uint8_t*>(c->alloc(perm.size()));
    if (!c->nperm) {
      set_error(alloc_failure("geometry"));
      return -3;
    }
    if (h2d(c, c->nperm, perm.data(), perm.size(), s)) return -3;
  }
  std::vector<int32_t> nb32(nz), bt32(nz), al32(nz);
  for (int64_t z = 0; z < nz; ++z) {
    nb32[z] = (int32_t)c->topo.nbf[z];
    bt32[z] = (int32_t)std::max<int64_t>(std::min<int64_t>(c->topo.bctype[z], INT_MAX), INT_MIN);
    al32[z] = (int32_t)std::max<int64_t>(c->topo.alnmt[z], 0);
  }
  const int nb = (int)((nele + GEO_THREADS - 1) / GEO_THREADS);
  double *dx = nullptr, *dk = nullptr, *dmax = nullptr;
  int32_t *dnb = nullptr, *dbt = nullptr, *dal = nullptr;
  int* dbad = nullptr;
  int* dbadf = nullptr;
  int rc = 0;
  auto free_all = [&]() {
    (void)hipStreamSynchronize(s);  // nothing in flight reads the work buffers any more
    for (void* p : {(void*)dx, (void*)dk, (void*)dmax, (void*)dnb, (void*)dbt, (void*)dal,
                    (void*)dbad, (void*)dbadf})
      dev_free(p);
  };
  auto step = [&](int r) {
    if (r && !rc) rc = r;
    return rc == 0;
  };
  const size_t nx = (size_t)nele * D * c->nloc;
  const size_t per_elem = (size_t)D * c->nloc;
  auto copy_parts = [&]() {
    for (int q = 0; q < nparts; ++q) {
      const size_t n = (size_t)(part_off[q + 1] - part_off[q]) * per_elem;
      if (n && h2d(c, dx + (size_t)part_off[q] * per_elem, x_parts[q], sizeof(double) * n, s))
        return -3;
    }
    return 0;
  };
  auto take = [&](auto** d, size_t bytes) {
    *d = static_cast<std::remove_pointer_t<decltype(d)>>(dev_alloc(bytes));
    return *d ? 0 : -3;
  };
  if (step(take(&dx, sizeof(double) * nx)) && step(copy_parts()) &&
      step(to_dev(c, c->kappa, &dk, s)) && step(to_dev(c, nb32, &dnb, s)) &&
      step(to_dev(c, bt32, &dbt, s)) && step(to_dev(c, al32, &dal, s)) && step(take(&dmax, sizeof(double) * nb)) &&
      step(take(&dbad, sizeof(int))) && step(take(&dbadf, sizeof(int)))) {
    const int big = INT_MAX;
    if (step(h2d(c, dbad, &big, sizeof(int), s) ? -3 : 0) &&
        step(h2d(c, dbadf, &big, sizeof(int), s) ? -3 : 0)) {
      if (D == 3)
        hipLaunchKernelGGL(k_geometry<3>, dim3(nb), dim3(GEO_THREADS), 0, s, (int)nele, dx, dk,
                           dnb, dbt, dal, t, g.G, g.cf, g.wf, g.pen, c->nbface, c->fcode,
                           c->detJ, g.cnb, dmax, dbad, dbadf);
      else
        hipLaunchKernelGGL(k_geometry<2>, dim3(nb), dim3(GEO_THREADS), 0, s, (int)nele, dx, dk,
                           dnb, dbt, dal, t, g.G, g.cf, g.wf, g.pen, c->nbface, c->fcode,
                           c->detJ, g.cnb, dmax, dbad, dbadf);
      std::vector<double> bm(nb);
      int hbad = INT_MAX, hbadf = INT_MAX;
      if (step(hipGetLastError() == hipSuccess ? 0 : -3) &&
          step(d2h(bm.data(), dmax, sizeof(double) * nb, s) ? -3 : 0) &&
          step(d2h(&hbad, dbad, sizeof(int), s) ? -3 : 0) &&
          step(d2h(&hbadf, dbadf, sizeof(int), s) ? -3 : 0)) {
        double m = 0.0;
        for (double v : bm) m = std::max(m, v);
        *max_dev = m;
        *bad_elem = hbad == INT_MAX ? -1 : hbad;
        // single-pass kernels need every interior face's frame
        c->frames_ok = hbadf == INT_MAX;
      }
    }
  }
  if (rc) set_error("geometry on the device failed: " + std::string(hipGetErrorString(hipGetLastError())));
  free_all();
  return rc;
}

}  // namespace dgmi
