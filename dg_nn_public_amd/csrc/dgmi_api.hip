// extern "C" entry points of libdgmi.so (declared in include/dgmi.h), context set-up,
// multigrid set-up, the V-cycle preconditioner and the GMRES driver.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <list>
#include <numeric>

#include "dgmi.h"
#include "dgmi_internal.h"

namespace dgmi {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

#define DGMI_CHECK(cond, msg)   \
  do {                          \
    if (!(cond)) {              \
      ::dgmi::set_error(msg);   \
      return -2;                \
    }                           \
  } while (0)
#define DGMI_TRY(call)            \
  do {                            \
    int rc_ = (call);             \
    if (rc_ != 0) return rc_;     \
  } while (0)

template <typename T, typename U>
static T* upload(dgmi_ctx* c, const std::vector<U>& h) {
  std::vector<T> t(h.begin(), h.end());
  T* d = static_cast<T*>(c->alloc(sizeof(T) * std::max<size_t>(t.size(), 1)));
  if (!d) return nullptr;
  if (!t.empty() && hipMemcpy(d, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  return d;
}

// kernel-friendly table layout (see DevGeom)
static std::vector<double> table_block(const ElementTables& el, size_t& ov, size_t& ow,
                                       size_t& of, size_t& os) {
  const int D = el.D, NL = el.nloc, NG = el.ngi, SG = el.sngi, NF = el.nface;
  std::vector<double> t;
  ov = t.size();
  for (int g = 0; g < NG; ++g)
    for (int d = 0; d < D; ++d)
      for (int n = 0; n < NL; ++n) t.push_back(el.nlx[(d * NL + n) * NG + g]);
  ow = t.size();
  for (int g = 0; g < NG; ++g) t.push_back(el.weight[g]);
  while (t.size() % 4) t.push_back(0.0);
  of = t.size();
  for (int f = 0; f < NF; ++f)
    for (int g = 0; g < SG; ++g) {
      for (int n = 0; n < NL; ++n) t.push_back(el.sn[(f * NL + n) * SG + g]);
      for (int d = 0; d < D; ++d)
        for (int n = 0; n < NL; ++n) t.push_back(el.snlx[((f * D + d) * NL + n) * SG + g]);
    }
  os = t.size();
  for (int g = 0; g < SG; ++g) t.push_back(el.sweight[g]);
  return t;
}

template <typename T>
static int upload_geom(dgmi_ctx* c, DevGeom<T>& g, const std::vector<double>& tab, size_t ov,
                       size_t ow, size_t of, size_t os, const Geometry& geo) {
  g.tab = upload<T>(c, tab);
  g.off_vtab = ov;
  g.off_wgt = ow;
  g.off_ftab = of;
  g.off_sw = os;
  g.G = upload<T>(c, geo.G);
  g.cf = upload<T>(c, geo.cf);
  g.wf = upload<T>(c, geo.wf);
  g.pen = upload<T>(c, geo.pen);
  g.trace = static_cast<typename Vec2<T>::type*>(
      c->alloc(sizeof(typename Vec2<T>::type) * c->nface * c->sngi * c->nele));
  DGMI_CHECK(g.tab && g.G && g.cf && g.wf && g.pen && g.trace, "device allocation failed");
  return 0;
}

// reference matrices for the block diagonal: Sref_jk, M_f, C_fd (see k_bdiag)
static std::vector<double> bdiag_refmats(const ElementTables& el) {
  const int D = el.D, NL = el.nloc, NG = el.ngi, SG = el.sngi, NF = el.nface, NN = NL * NL;
  // Sref (D*D*NN) | M_f (NF*NN) | C_fd (NF*D*NN) | Mref (NN), Mref[n][m] = sum_g w n_n n_m
  std::vector<double> r((size_t)D * D * NN + (size_t)NF * NN + (size_t)NF * D * NN + NN, 0.0);
  double* Mref = r.data() + (size_t)D * D * NN + (size_t)NF * NN + (size_t)NF * D * NN;
  for (int n = 0; n < NL; ++n)
    for (int m = 0; m < NL; ++m) {
      double s = 0.0;
      for (int g = 0; g < NG; ++g) s += el.weight[g] * el.n[n * NG + g] * el.n[m * NG + g];
      Mref[n * NL + m] = s;
    }
  for (int j = 0; j < D; ++j)
    for (int k = 0; k < D; ++k)
      for (int n = 0; n < NL; ++n)
        for (int m = 0; m < NL; ++m) {
          double s = 0.0;
          for (int g = 0; g < NG; ++g)
            s += el.weight[g] * el.nlx[(j * NL + n) * NG + g] * el.nlx[(k * NL + m) * NG + g];
          r[(size_t)(j * D + k) * NN + n * NL + m] = s;
        }
  double* M = r.data() + (size_t)D * D * NN;
  double* C = M + (size_t)NF * NN;
  for (int f = 0; f < NF; ++f)
    for (int n = 0; n < NL; ++n)
      for (int m = 0; m < NL; ++m) {
        double s = 0.0;
        for (int g = 0; g < SG; ++g)
          s += el.sweight[g] * el.sn[(f * NL + n) * SG + g] * el.sn[(f * NL + m) * SG + g];
        M[(size_t)f * NN + n * NL + m] = s;
        for (int d = 0; d < D; ++d) {
          double t = 0.0;
          for (int g = 0; g < SG; ++g)
            t += el.sweight[g] * el.sn[(f * NL + n) * SG + g] *
                 el.snlx[((f * D + d) * NL + m) * SG + g];
          C[(size_t)(f * D + d) * NN + n * NL + m] = t;
        }
      }
  return r;
}

// Gauss-Jordan inverse with partial pivoting of one n x n block (row-major), in place
static bool invert_block(double* a, int n) {
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) return false;
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[p * n + k], a[c * n + k]);
        std::swap(inv[p * n + k], inv[c * n + k]);
      }
    const double d = 1.0 / a[c * n + c];
    for (int k = 0; k < n; ++k) {
      a[c * n + k] *= d;
      inv[c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r)
      if (r != c) {
        const double f = a[r * n + c];
        if (f != 0.0)
          for (int k = 0; k < n; ++k) {
            a[r * n + k] -= f * a[c * n + k];
            inv[r * n + k] -= f * inv[c * n + k];
          }
      }
  }
  std::copy(inv.begin(), inv.end(), a);
  return true;
}

struct Scratch {
  double* bdiag = nullptr;
  double* refm = nullptr;
  double* vec64[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float* vec32[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* cg64[2] = {nullptr, nullptr};
  float* cg32[2] = {nullptr, nullptr};
};
static Scratch& scratch(dgmi_ctx* c);

}  // namespace dgmi

using namespace dgmi;

void* dgmi_ctx::alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  allocs.push_back(p);
  dev_bytes += bytes;
  return p;
}

namespace dgmi {
static std::list<std::pair<dgmi_ctx*, Scratch>> g_scratch;  // stable references
static Scratch& scratch(dgmi_ctx* c) {
  for (auto& p : g_scratch)
    if (p.first == c) return p.second;
  g_scratch.emplace_back(c, Scratch{});
  return g_scratch.back().second;
}
template <typename T> static T* vbuf(dgmi_ctx* c, int i);
template <> double* vbuf<double>(dgmi_ctx* c, int i) {
  Scratch& s = scratch(c);
  if (!s.vec64[i]) s.vec64[i] = static_cast<double*>(c->alloc(sizeof(double) * c->ndof));
  return s.vec64[i];
}
template <> float* vbuf<float>(dgmi_ctx* c, int i) {
  Scratch& s = scratch(c);
  if (!s.vec32[i]) s.vec32[i] = static_cast<float*>(c->alloc(sizeof(float) * c->ndof));
  return s.vec32[i];
}
// slot 4: the GMRES correction vector
template <typename T> static T* corr_buf(dgmi_ctx* c) { return vbuf<T>(c, 4); }
template <typename T> static T* cgbuf(dgmi_ctx* c, int i);
template <> double* cgbuf<double>(dgmi_ctx* c, int i) {
  Scratch& s = scratch(c);
  if (!s.cg64[i]) s.cg64[i] = static_cast<double*>(c->alloc(sizeof(double) * c->mg.cg_nonods));
  return s.cg64[i];
}
template <> float* cgbuf<float>(dgmi_ctx* c, int i) {
  Scratch& s = scratch(c);
  if (!s.cg32[i]) s.cg32[i] = static_cast<float*>(c->alloc(sizeof(float) * c->mg.cg_nonods));
  return s.cg32[i];
}
template <typename T> static DevGeom<T>& geom(dgmi_ctx* c);
template <> DevGeom<double>& geom<double>(dgmi_ctx* c) { return c->g64; }
template <> DevGeom<float>& geom<float>(dgmi_ctx* c) { return c->g32; }

static int compute_bdiag(dgmi_ctx* c, hipStream_t s) {
  Scratch& sc = scratch(c);
  if (!sc.refm) {
    sc.refm = upload<double>(c, bdiag_refmats(c->el));
    DGMI_CHECK(sc.refm, "device allocation failed");
  }
  if (!sc.bdiag) {
    sc.bdiag = static_cast<double*>(c->alloc(sizeof(double) * c->nele * c->nloc * c->nloc));
    DGMI_CHECK(sc.bdiag, "device allocation failed");
  }
  return launch_bdiag(c, sc.refm, sc.bdiag, s);
}

// _um_left_precond: x = M^-1 b, one V-cycle from x = 0 (z02/volume_mf_diff.py:1934-2029)
template <typename T>
static int precond(dgmi_ctx* c, const T* b, T* x, hipStream_t s) {
  DGMI_CHECK(c->smoother_variant >= 0, "smoother not set up (dgmi_smoother_setup)");
  DGMI_CHECK(c->mg.ready, "multigrid not set up (dgmi_mg_setup)");
  DevGeom<T>& g = geom<T>(c);
  T* r = vbuf<T>(c, 3);
  T* rc = cgbuf<T>(c, 0);
  T* ec = cgbuf<T>(c, 1);
  // pre-smoothing: the first step starts from x = 0, where the residual is b
  DGMI_TRY(launch_smooth_zero<T>(c, g, x, b, s));
  for (int it = 1; it < 3; ++it) DGMI_TRY(launch_smooth<T>(c, g, x, b, nullptr, s));
  DGMI_TRY(launch_apply<T>(c, g, x, b, r, s));
  DGMI_TRY(launch_restrict<T>(c, r, rc, s));
  DGMI_TRY(launch_sfc_vcycle<T>(c, rc, ec, s));
  DGMI_TRY(launch_prolong_add<T>(c, ec, x, s));
  // post-smoothing (the reference's loop runs pre_smooth_its times, :2021)
  for (int it = 0; it < 3; ++it) DGMI_TRY(launch_smooth<T>(c, g, x, b, nullptr, s));
  return 0;
}

// Solver workspace: Krylov basis, Hessenberg, rotations (device), pinned host scalars.
struct Krylov {
  int m = 0;
  void* V = nullptr;  // (m+1) * ndof of T
  size_t vbytes = 0;
  double* H = nullptr;   // (m+1) x m column-major, ldh = m+1
  double* cs = nullptr;
  double* sn = nullptr;
  double* g = nullptr;
  double* y = nullptr;
  double* red = nullptr;  // reduction scalars
  double* host = nullptr;  // pinned
};
static std::list<std::pair<dgmi_ctx*, Krylov>> g_kry;
static Krylov* krylov(dgmi_ctx* c, int m, size_t elem) {
  Krylov* k = nullptr;
  for (auto& p : g_kry)
    if (p.first == c) k = &p.second;
  if (!k) {
    g_kry.emplace_back(c, Krylov{});
    k = &g_kry.back().second;
  }
  const size_t need = (size_t)(m + 1) * c->ndof * elem;
  if (k->m < m || k->vbytes < need) {
    k->V = c->alloc(need);
    k->vbytes = need;
    const int ldh = m + 1;
    k->H = static_cast<double*>(c->alloc(sizeof(double) * ldh * m));
    k->cs = static_cast<double*>(c->alloc(sizeof(double) * (m + 1)));
    k->sn = static_cast<double*>(c->alloc(sizeof(double) * (m + 1)));
    k->g = static_cast<double*>(c->alloc(sizeof(double) * (m + 2)));
    k->y = static_cast<double*>(c->alloc(sizeof(double) * (m + 1)));
    k->m = m;
  }
  if (!k->red) k->red = static_cast<double*>(c->alloc(sizeof(double) * 8));
  if (!k->host && hipHostMalloc((void**)&k->host, sizeof(double) * 8) != hipSuccess) return nullptr;
  if (!k->V || !k->H || !k->g || !k->red || !k->host) return nullptr;
  return k;
}

// One GMRES cycle on A d = r0 from d = 0, left-preconditioned, modified Gram-Schmidt,
// Givens rotations (Saad, Alg. 6.9).  Stops at m iterations or when the preconditioned
// residual estimate |g_{j+1}| <= eps * beta.  Returns iterations used in *its.
template <typename T>
static int gmres_cycle(dgmi_ctx* c, const T* r0, T* d, int m, double eps, int* its,
                       hipStream_t s) {
  Krylov* k = krylov(c, m, sizeof(T));
  DGMI_CHECK(k, "device allocation failed (Krylov workspace)");
  const int64_t n = c->ndof;
  const int ldh = m + 1;
  T* V = static_cast<T*>(k->V);
  T* w = vbuf<T>(c, 2);
  T* t = vbuf<T>(c, 1);
  auto Vj = [&](int j) { return V + (size_t)j * n; };
  DGMI_TRY(precond<T>(c, r0, Vj(0), s));
  DGMI_TRY(launch_dot<T>(n, Vj(0), Vj(0), k->red, 1, s));  // red[1] = beta
  DGMI_TRY(kry::scale_dev<T>(n, k->red + 1, Vj(0), Vj(0), s));
  DGMI_TRY(kry::set_g(k->g, m + 2, k->red, s));
  DGMI_HIP(hipMemcpyAsync(k->host, k->red + 1, sizeof(double), hipMemcpyDeviceToHost, s));
  DGMI_HIP(hipStreamSynchronize(s));
  const double beta = k->host[0];
  DGMI_CHECK(std::isfinite(beta), "non-finite preconditioned residual");
  int j = 0;
  for (; j < m; ++j) {
    DGMI_TRY(launch_apply_pos<T>(c, geom<T>(c), Vj(j), t, s));  // t = A v_j
    DGMI_TRY(precond<T>(c, t, w, s));                           // w = M^-1 A v_j
    double* hcol = k->H + (size_t)j * ldh;
    if (c->gmres_orth == 0) {
      // modified Gram-Schmidt, as the reference (z02/volume_mf_diff.py:1813-1816)
      for (int i = 0; i <= j; ++i) {
        DGMI_TRY(launch_dot<T>(n, w, Vj(i), hcol + i, 0, s));
        DGMI_TRY(kry::axpy_dev<T>(n, hcol + i, -1.0, Vj(i), w, s));
      }
    } else {
      // classical Gram-Schmidt with one re-orthogonalisation (CGS2): two fused passes,
      // the Hessenberg column is the sum of both passes' coefficients
      DGMI_TRY(kry::cgs_pass<T>(n, j + 1, V, n, w, hcol, 0, s));
      DGMI_TRY(kry::cgs_pass<T>(n, j + 1, V, n, w, hcol, 1, s));
    }
    DGMI_TRY(launch_dot<T>(n, w, w, k->red, 1, s));
    DGMI_TRY(kry::store_h(k->red + 1, hcol + j + 1, 0, s));
    DGMI_TRY(kry::scale_dev<T>(n, k->red + 1, w, Vj(j + 1), s));
    DGMI_TRY(kry::givens(j, ldh, k->H, k->cs, k->sn, k->g, k->red + 2, s));
    DGMI_HIP(hipMemcpyAsync(k->host + 1, k->red + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    DGMI_HIP(hipStreamSynchronize(s));
    const double est = k->host[1];
    DGMI_CHECK(std::isfinite(est), "non-finite Arnoldi residual estimate");
    if (est <= eps * beta) {
      ++j;
      break;
    }
  }
  const int kk = std::min(j, m);
  DGMI_TRY(kry::backsolve(kk, ldh, k->H, k->g, k->y, s));
  DGMI_TRY(kry::fill<T>(n, d, T(0), s));
  DGMI_TRY(kry::combine<T>(n, kk, V, n, k->y, d, s));
  *its = kk;
  return 0;
}

}  // namespace dgmi

// ========================================================================================
extern "C" {

const char* dgmi_last_error(void) { return g_err.c_str(); }

int dgmi_ctx_create(const dgmi_mesh_desc* d, dgmi_ctx** out) {
  DGMI_CHECK(d && out, "null argument");
  DGMI_CHECK(d->D == 2 || d->D == 3, "D must be 2 or 3");
  DGMI_CHECK(d->nele > 0 && d->nverts > 0, "empty mesh");
  DGMI_CHECK(d->x_ref_in && d->cells && d->verts && d->kappa, "null mesh array");
  DGMI_CHECK(d->nele <= 4000000, "nele above the 32-bit index range of the kernels (4e6)");
  auto* c = new dgmi_ctx();
  c->D = d->D;
  c->nele = d->nele;
  c->nverts = d->nverts;
  c->eta_e = d->eta_e > 0 ? d->eta_e : 36.0;
  c->el = make_element(c->D);
  c->nloc = c->el.nloc;
  c->ngi = c->el.ngi;
  c->sngi = c->el.sngi;
  c->nface = c->el.nface;
  c->ng = c->D == 3 ? 6 : 3;
  c->ndof = c->nele * c->nloc;
  c->cells.assign(d->cells, d->cells + c->nele * (c->D + 1));
  c->verts.assign(d->verts, d->verts + c->nverts * c->D);
  c->kappa.assign(d->kappa, d->kappa + c->nele);
  for (int64_t v : c->cells)
    if (v < 0 || v >= c->nverts) {
      delete c;
      set_error("cell vertex id out of range");
      return -2;
    }
  c->topo = build_topology(c->cells.data(), c->nele, c->D);
  if (d->bctype) {
    for (int64_t z = 0; z < c->nele * c->nface; ++z) {
      const int64_t t = d->bctype[z];
      if (t < 0 && c->topo.nbf[z] < 0) {
        delete c;
        set_error("face marked interior has no neighbour");
        return -2;
      }
      c->topo.bctype[z] = t;
    }
  }
  Geometry geo;
  std::string err;
  DGMI_CHECK(d->face_kappa == 0 || d->face_kappa == 1, "face_kappa must be 0 or 1");
  if (!build_geometry(c->el, d->x_ref_in, c->kappa.data(), c->topo, c->nele, c->eta_e,
                      d->face_kappa, geo, err)) {
    delete c;
    set_error(err);
    return -2;
  }
  if (geo.max_affine_dev > 1e-8) {
    delete c;
    set_error("curved (non-affine) P3 elements are not supported: max node deviation " +
              std::to_string(geo.max_affine_dev));
    return -2;
  }
  size_t ov, ow, of, os;
  const std::vector<double> tab = table_block(c->el, ov, ow, of, os);
  int rc = upload_geom<double>(c, c->g64, tab, ov, ow, of, os, geo);
  if (!rc) rc = upload_geom<float>(c, c->g32, tab, ov, ow, of, os, geo);
  if (!rc) {
    c->nbface = upload<int32_t>(c, geo.nbface);
    c->fcode = upload<int32_t>(c, geo.fcode);
    std::vector<int32_t> al(c->el.gi_align.begin(), c->el.gi_align.end());
    c->aln = upload<int32_t>(c, al);
    c->detJ = upload<double>(c, geo.detJ);
    if (!c->nbface || !c->fcode || !c->aln || !c->detJ) {
      set_error("device allocation failed");
      rc = -3;
    }
  }
  if (rc) {
    dgmi_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int dgmi_ctx_destroy(dgmi_ctx* c) {
  if (!c) return 0;
  (void)hipDeviceSynchronize();
  for (void* p : c->allocs) (void)hipFree(p);
  for (auto it = g_scratch.begin(); it != g_scratch.end(); ++it)
    if (it->first == c) {
      g_scratch.erase(it);
      break;
    }
  for (auto it = g_kry.begin(); it != g_kry.end(); ++it)
    if (it->first == c) {
      if (it->second.host) (void)hipHostFree(it->second.host);
      g_kry.erase(it);
      break;
    }
  delete c;
  return 0;
}

int dgmi_ctx_info(const dgmi_ctx* c, int64_t* info) {
  DGMI_CHECK(c && info, "null argument");
  info[0] = c->nele;
  info[1] = c->nloc;
  info[2] = c->ndof;
  info[3] = c->nface;
  info[4] = c->ngi;
  info[5] = c->sngi;
  info[6] = c->nverts;
  info[7] = (int64_t)c->dev_bytes;
  return 0;
}

int dgmi_topology(const dgmi_ctx* c, int64_t* nbf, int64_t* alnmt, int64_t* bctype) {
  DGMI_CHECK(c, "null context");
  const size_t n = c->topo.nbf.size();
  if (nbf) std::memcpy(nbf, c->topo.nbf.data(), n * sizeof(int64_t));
  if (alnmt) std::memcpy(alnmt, c->topo.alnmt.data(), n * sizeof(int64_t));
  if (bctype) std::memcpy(bctype, c->topo.bctype.data(), n * sizeof(int64_t));
  return 0;
}

int dgmi_element_tables(const dgmi_ctx* c, double* n, double* nlx, double* weight, double* sn,
                        double* snlx, double* sweight, int64_t* gi_align) {
  DGMI_CHECK(c, "null context");
  const ElementTables& e = c->el;
  if (n) std::copy(e.n.begin(), e.n.end(), n);
  if (nlx) std::copy(e.nlx.begin(), e.nlx.end(), nlx);
  if (weight) std::copy(e.weight.begin(), e.weight.end(), weight);
  if (sn) std::copy(e.sn.begin(), e.sn.end(), sn);
  if (snlx) std::copy(e.snlx.begin(), e.snlx.end(), snlx);
  if (sweight) std::copy(e.sweight.begin(), e.sweight.end(), sweight);
  if (gi_align) std::copy(e.gi_align.begin(), e.gi_align.end(), gi_align);
  return 0;
}

static int check_align(const void* p) {
  DGMI_CHECK(((uintptr_t)p & 15) == 0, "DOF vectors must be 16-byte aligned");
  return 0;
}

int dgmi_apply(dgmi_ctx* c, int dtype, const void* x, const void* b, void* r, void* stream) {
  DGMI_CHECK(c && x && r, "null argument");
  DGMI_TRY(check_align(x));
  DGMI_TRY(check_align(r));
  if (b) DGMI_TRY(check_align(b));
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64)
    return launch_apply<double>(c, c->g64, (const double*)x, (const double*)b, (double*)r, s);
  if (dtype == DGMI_F32)
    return launch_apply<float>(c, c->g32, (const float*)x, (const float*)b, (float*)r, s);
  set_error("dtype not supported by dgmi_apply");
  return -2;
}

int dgmi_rhs(dgmi_ctx* c, const double* f, const double* u_bc, double* b, void* stream) {
  DGMI_CHECK(c && f && u_bc && b, "null argument");
  hipStream_t s = (hipStream_t)stream;
  Scratch& sc = scratch(c);
  if (!sc.refm) {
    sc.refm = upload<double>(c, bdiag_refmats(c->el));
    DGMI_CHECK(sc.refm, "device allocation failed");
  }
  return launch_rhs(c, sc.refm, c->detJ, f, u_bc, b, s);
}

int dgmi_bdiag(dgmi_ctx* c, double* bdiag, double* diag, void* stream) {
  DGMI_CHECK(c, "null context");
  hipStream_t s = (hipStream_t)stream;
  DGMI_TRY(compute_bdiag(c, s));
  Scratch& sc = scratch(c);
  if (bdiag)
    DGMI_HIP(hipMemcpyAsync(bdiag, sc.bdiag, sizeof(double) * c->nele * c->nloc * c->nloc,
                            hipMemcpyDeviceToDevice, s));
  if (diag) DGMI_TRY(launch_extract_diag(c, sc.bdiag, diag, s));
  return 0;
}

int dgmi_smoother_setup(dgmi_ctx* c, int variant, double omega, void* stream) {
  DGMI_CHECK(c, "null context");
  DGMI_CHECK(variant == DGMI_SMOOTH_JACOBI3 || variant == DGMI_SMOOTH_DIRECT,
             "unknown smoother variant");
  hipStream_t s = (hipStream_t)stream;
  DGMI_TRY(compute_bdiag(c, s));
  const int64_t nS = c->nele * c->nloc * c->nloc;
  if (!c->g64.S) c->g64.S = static_cast<double*>(c->alloc(sizeof(double) * nS));
  if (!c->g32.S) c->g32.S = static_cast<float*>(c->alloc(sizeof(float) * nS));
  DGMI_CHECK(c->g64.S && c->g32.S, "device allocation failed");
  Scratch& sc = scratch(c);
  if (variant == DGMI_SMOOTH_JACOBI3) {
    DGMI_TRY(launch_build_S_jacobi<double>(c, sc.bdiag, omega, c->g64.S, s));
    DGMI_TRY(launch_build_S_jacobi<float>(c, sc.bdiag, omega, c->g32.S, s));
  } else {
    const int NL = c->nloc, NN = NL * NL;
    std::vector<double> B((size_t)nS), St((size_t)nS);
    std::vector<float> Sf((size_t)nS);
    DGMI_HIP(hipMemcpyAsync(B.data(), sc.bdiag, sizeof(double) * nS, hipMemcpyDeviceToHost, s));
    DGMI_HIP(hipStreamSynchronize(s));
    // structure-of-vectors layout of the kernels (width W: 16 bytes, or 8 when the row
    // length is not a multiple of 16 bytes)
    auto width = [&](size_t bytes) { return (NL % (16 / bytes) == 0) ? 16 / bytes : 8 / bytes; };
    const int W64 = (int)width(8), W32 = (int)width(4);
    const int64_t ne = c->nele;
    for (int64_t e = 0; e < ne; ++e) {
      double* blk = B.data() + e * NN;
      DGMI_CHECK(invert_block(blk, NL), "singular diagonal block " + std::to_string(e));
      for (int n = 0; n < NL; ++n)
        for (int m = 0; m < NL; ++m) {
          const double v = omega * blk[n * NL + m];
          St[((size_t)(n * (NL / W64) + m / W64) * ne + e) * W64 + m % W64] = v;
          Sf[((size_t)(n * (NL / W32) + m / W32) * ne + e) * W32 + m % W32] = (float)v;
        }
    }
    DGMI_HIP(hipMemcpy(c->g64.S, St.data(), sizeof(double) * nS, hipMemcpyHostToDevice));
    DGMI_HIP(hipMemcpy(c->g32.S, Sf.data(), sizeof(float) * nS, hipMemcpyHostToDevice));
  }
  c->smoother_variant = variant;
  c->omega = omega;
  return 0;
}

int dgmi_smooth(dgmi_ctx* c, int dtype, void* x, const void* b, void* r_out, void* stream) {
  DGMI_CHECK(c && x && b, "null argument");
  DGMI_CHECK(c->smoother_variant >= 0, "smoother not set up (dgmi_smoother_setup)");
  DGMI_TRY(check_align(x));
  DGMI_TRY(check_align(b));
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64)
    return launch_smooth<double>(c, c->g64, (double*)x, (const double*)b, (double*)r_out, s);
  if (dtype == DGMI_F32)
    return launch_smooth<float>(c, c->g32, (float*)x, (const float*)b, (float*)r_out, s);
  set_error("dtype not supported by dgmi_smooth");
  return -2;
}

int dgmi_mg_setup(dgmi_ctx* c, void* stream) {
  DGMI_CHECK(c, "null context");
  hipStream_t s = (hipStream_t)stream;
  MGDev& mg = c->mg;
  const int D = c->D, NB = D + 1;
  p1cg_sparsity(c->cells.data(), c->nele, D, c->nverts, mg.fina, mg.cola);
  mg.cg_nonods = c->nverts;
  mg.ncola = (int64_t)mg.cola.size();
  std::vector<int32_t> whichc;
  mg.ncolor = color2(mg.fina, mg.cola, whichc);
  hilbert_order(c->verts.data(), c->nverts, D, mg.sfc);
  const int64_t ncg = mg.cg_nonods;
  // restriction lists by SFC position; multiplicities
  std::vector<int32_t> mult(ncg, 0);
  for (int64_t v : c->cells) mult[v]++;
  std::vector<int32_t> rptr(ncg + 1, 0), rent(c->cells.size()), cellpos(c->cells.size());
  for (int64_t i = 0; i < ncg; ++i) rptr[mg.sfc[i]] = mult[i];  // position sfc[i]-1, +1 shift
  for (int64_t p = 0; p < ncg; ++p) rptr[p + 1] += rptr[p];
  std::vector<int32_t> fillp(rptr.begin(), rptr.end() - 1);
  for (int64_t q = 0; q < (int64_t)c->cells.size(); ++q) {
    const int64_t p = mg.sfc[c->cells[q]] - 1;
    rent[fillp[p]++] = (int32_t)q;
    cellpos[q] = (int32_t)p;
  }
  std::vector<double> rw(ncg);
  std::vector<int32_t> whichc_sfc(ncg);
  for (int64_t i = 0; i < ncg; ++i) {
    DGMI_CHECK(mult[i] > 0, "P1 vertex not used by any element");
    rw[mg.sfc[i] - 1] = 1.0 / mult[i];
    whichc_sfc[mg.sfc[i] - 1] = whichc[i];
  }
  mg.rptr = upload<int32_t>(c, rptr);
  mg.rent = upload<int32_t>(c, rent);
  mg.cellpos = upload<int32_t>(c, cellpos);
  mg.rw64 = upload<double>(c, rw);
  mg.rw32 = upload<float>(c, rw);
  int32_t* whichc_d = upload<int32_t>(c, whichc_sfc);
  DGMI_CHECK(mg.rptr && mg.rent && mg.cellpos && mg.rw64 && mg.rw32 && whichc_d,
             "device allocation failed");
  // coloured probing of RAR (z02/volume_mf_diff.py:1696-1754): entries grouped by the
  // colour of their column; each row holds at most one entry per colour
  std::vector<int64_t> cnt(mg.ncolor + 1, 0);
  for (int64_t q = 0; q < mg.ncola; ++q) cnt[whichc[mg.cola[q]]]++;
  std::vector<int64_t> off(mg.ncolor + 2, 0);
  for (int cc = 1; cc <= mg.ncolor; ++cc) off[cc + 1] = off[cc] + cnt[cc];
  std::vector<int32_t> spidx(mg.ncola), rowpos(mg.ncola);
  {
    std::vector<int64_t> f2(off.begin(), off.end());
    for (int64_t i = 0; i < ncg; ++i)
      for (int64_t q = mg.fina[i]; q < mg.fina[i + 1]; ++q) {
        const int cc = whichc[mg.cola[q]];
        spidx[f2[cc]] = (int32_t)q;
        rowpos[f2[cc]] = (int32_t)(mg.sfc[i] - 1);
        f2[cc]++;
      }
  }
  int32_t* spidx_d = upload<int32_t>(c, spidx);
  int32_t* rowpos_d = upload<int32_t>(c, rowpos);
  double* vals_d = static_cast<double*>(c->alloc(sizeof(double) * mg.ncola));
  double* probe = vbuf<double>(c, 0);
  double* aprobe = vbuf<double>(c, 1);
  double* rsfc = cgbuf<double>(c, 0);
  DGMI_CHECK(spidx_d && rowpos_d && vals_d && probe && aprobe && rsfc, "device allocation failed");
  for (int cc = 1; cc <= mg.ncolor; ++cc) {
    DGMI_TRY(launch_probe_mask(c, whichc_d, cc, probe, s));
    DGMI_TRY(launch_apply<double>(c, c->g64, probe, nullptr, aprobe, s));  // -A R m
    DGMI_TRY(launch_restrict<double>(c, aprobe, rsfc, s));
    DGMI_TRY(launch_probe_collect(c, rsfc, spidx_d + off[cc], rowpos_d + off[cc],
                                  off[cc + 1] - off[cc], vals_d, s));
  }
  mg.rar.assign(mg.ncola, 0.0);
  DGMI_HIP(hipMemcpyAsync(mg.rar.data(), vals_d, sizeof(double) * mg.ncola,
                          hipMemcpyDeviceToHost, s));
  DGMI_HIP(hipStreamSynchronize(s));
  // SFC levels (z00-fortran-src/compact_fcode.f90:7-225)
  SfcLevelsHost lh;
  build_sfc_levels(mg.fina, mg.cola, mg.rar, mg.sfc, lh);
  SfcLevelsDev& lv = mg.lv;
  lv.nlevel = lh.nlevel;
  lv.n = lh.n;
  const int L = lh.nlevel;
  lv.rowptr.resize(L);
  lv.col.resize(L);
  lv.val64.resize(L);
  lv.diag64.resize(L);
  lv.val32.resize(L);
  lv.diag32.resize(L);
  lv.r.resize(L);
  lv.e.resize(L);
  lv.e2.resize(L);
  for (int l = 0; l < L; ++l) {
    lv.rowptr[l] = upload<int32_t>(c, lh.rowptr[l]);
    lv.col[l] = upload<int32_t>(c, lh.col[l]);
    lv.val64[l] = upload<double>(c, lh.val[l]);
    lv.diag64[l] = upload<double>(c, lh.diag[l]);
    lv.val32[l] = upload<float>(c, lh.val[l]);
    lv.diag32[l] = upload<float>(c, lh.diag[l]);
    lv.r[l] = c->alloc(sizeof(double) * lh.n[l]);
    lv.e[l] = c->alloc(sizeof(double) * lh.n[l]);
    lv.e2[l] = c->alloc(sizeof(double) * lh.n[l]);
    DGMI_CHECK(lv.rowptr[l] && lv.col[l] && lv.val64[l] && lv.diag64[l] && lv.val32[l] &&
                   lv.diag32[l] && lv.r[l] && lv.e[l] && lv.e2[l],
               "device allocation failed");
  }
  // levels with at most SMALL_N nodes run in the single-work-group kernel
  // (DGMI_SMALL_LEVEL_N overrides the threshold, e.g. to exercise the grid kernels on
  // small meshes in tests)
  // The coarse end of the hierarchy that fits in LDS (120 KiB of the 160 KiB per work-group)
  // runs in one work-group; larger levels run as grid kernels.
  lv.nnz.resize(L);
  for (int l = 0; l < L; ++l) lv.nnz[l] = (int64_t)lh.col[l].size();
  int l0 = choose_small_level(lv, lv.nnz, 120 * 1024);
  DGMI_CHECK(l0 >= 0, "too many SFC levels");
  if (const char* ev = std::getenv("DGMI_SMALL_LEVEL_N")) {
    const int64_t small_n = std::max<int64_t>(1, std::atoll(ev));
    for (int l = 0; l < L; ++l)
      if (lh.n[l] <= small_n) {
        l0 = std::max(l0, std::min(l, L - 1));
        break;
      }
  }
  mg.small_level = l0;
  DGMI_TRY(set_small_lds_attr<double>(lds_bytes_from<double>(lv, lv.nnz, l0)));
  DGMI_TRY(set_small_lds_attr<float>(lds_bytes_from<float>(lv, lv.nnz, l0)));
  DGMI_TRY(upload_level_descs(c));
  mg.ready = true;
  return 0;
}

int dgmi_mg_info(const dgmi_ctx* c, int64_t* info, int cap) {
  DGMI_CHECK(c && info, "null argument");
  DGMI_CHECK(c->mg.ready, "multigrid not set up");
  const int L = c->mg.lv.nlevel;
  DGMI_CHECK(cap >= 4 + L, "info buffer too small");
  info[0] = c->mg.cg_nonods;
  info[1] = c->mg.ncola;
  info[2] = c->mg.ncolor;
  info[3] = L;
  for (int l = 0; l < L; ++l) info[4 + l] = c->mg.lv.n[l];
  return 0;
}

int dgmi_mg_export(const dgmi_ctx* c, int64_t* fina, int64_t* cola, double* rar, int64_t* sfc) {
  DGMI_CHECK(c && c->mg.ready, "multigrid not set up");
  if (fina) std::copy(c->mg.fina.begin(), c->mg.fina.end(), fina);
  if (cola) std::copy(c->mg.cola.begin(), c->mg.cola.end(), cola);
  if (rar) std::copy(c->mg.rar.begin(), c->mg.rar.end(), rar);
  if (sfc) std::copy(c->mg.sfc.begin(), c->mg.sfc.end(), sfc);
  return 0;
}

int dgmi_restrict(dgmi_ctx* c, int dtype, const void* r_dg, void* r_sfc, void* stream) {
  DGMI_CHECK(c && r_dg && r_sfc, "null argument");
  DGMI_CHECK(c->mg.ready, "multigrid not set up (dgmi_mg_setup)");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64) return launch_restrict<double>(c, (const double*)r_dg, (double*)r_sfc, s);
  if (dtype == DGMI_F32) return launch_restrict<float>(c, (const float*)r_dg, (float*)r_sfc, s);
  set_error("dtype not supported");
  return -2;
}

int dgmi_prolong_add(dgmi_ctx* c, int dtype, const void* e_sfc, void* x_dg, void* stream) {
  DGMI_CHECK(c && e_sfc && x_dg, "null argument");
  DGMI_CHECK(c->mg.ready, "multigrid not set up (dgmi_mg_setup)");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64)
    return launch_prolong_add<double>(c, (const double*)e_sfc, (double*)x_dg, s);
  if (dtype == DGMI_F32) return launch_prolong_add<float>(c, (const float*)e_sfc, (float*)x_dg, s);
  set_error("dtype not supported");
  return -2;
}

int dgmi_sfc_vcycle(dgmi_ctx* c, int dtype, const void* r_sfc, void* e_sfc, void* stream) {
  DGMI_CHECK(c && r_sfc && e_sfc, "null argument");
  DGMI_CHECK(c->mg.ready, "multigrid not set up (dgmi_mg_setup)");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64)
    return launch_sfc_vcycle<double>(c, (const double*)r_sfc, (double*)e_sfc, s);
  if (dtype == DGMI_F32) return launch_sfc_vcycle<float>(c, (const float*)r_sfc, (float*)e_sfc, s);
  set_error("dtype not supported");
  return -2;
}

int dgmi_precond(dgmi_ctx* c, int dtype, const void* b, void* x, void* stream) {
  DGMI_CHECK(c && b && x, "null argument");
  DGMI_TRY(check_align(b));
  DGMI_TRY(check_align(x));
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64) return precond<double>(c, (const double*)b, (double*)x, s);
  if (dtype == DGMI_F32) return precond<float>(c, (const float*)b, (float*)x, s);
  set_error("dtype not supported");
  return -2;
}

int dgmi_gmres_solve(dgmi_ctx* c, int inner, const double* b, double* x, double rtol, int m,
                     int max_its, double* stats, void* stream) {
  DGMI_CHECK(c && b && x, "null argument");
  DGMI_CHECK(m >= 1 && m <= 200, "restart length m must be in [1, 200]");
  DGMI_CHECK(inner == DGMI_F64 || inner == DGMI_F32, "inner dtype not supported");
  DGMI_TRY(check_align(b));
  DGMI_TRY(check_align(x));
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = c->ndof;
  Krylov* k = krylov(c, m, inner == DGMI_F64 ? sizeof(double) : sizeof(float));
  DGMI_CHECK(k, "device allocation failed (Krylov workspace)");
  double* r64 = vbuf<double>(c, 0);
  DGMI_CHECK(r64, "device allocation failed");
  // ||b||
  DGMI_TRY(launch_dot<double>(n, b, b, k->red + 4, 1, s));
  DGMI_HIP(hipMemcpyAsync(k->host + 4, k->red + 5, sizeof(double), hipMemcpyDeviceToHost, s));
  DGMI_HIP(hipStreamSynchronize(s));
  const double bnorm = k->host[4];
  DGMI_CHECK(std::isfinite(bnorm), "non-finite right-hand side");
  const double tol = rtol * bnorm;
  int its = 0, outer = 0, cycles = 0;
  double rnorm = 0.0;
  for (;;) {
    DGMI_TRY(launch_apply<double>(c, c->g64, x, b, r64, s));  // FP64 true residual
    DGMI_TRY(launch_dot<double>(n, r64, r64, k->red + 4, 1, s));
    DGMI_HIP(hipMemcpyAsync(k->host + 4, k->red + 5, sizeof(double), hipMemcpyDeviceToHost, s));
    DGMI_HIP(hipStreamSynchronize(s));
    rnorm = k->host[4];
    DGMI_CHECK(std::isfinite(rnorm), "non-finite residual");
    if (rnorm <= tol || its >= max_its || bnorm == 0.0) break;
    const double eps = std::min(0.5, 0.5 * tol / rnorm);
    int used = 0;
    if (inner == DGMI_F64) {
      // the cycle uses scratch slots 1..3; the correction has its own slot
      double* dv = corr_buf<double>(c);
      DGMI_CHECK(dv, "device allocation failed");
      DGMI_TRY(gmres_cycle<double>(c, r64, dv, m, eps, &used, s));
      DGMI_TRY(kry::add_cast<double>(n, dv, x, nullptr, s));
    } else {
      float* r32 = vbuf<float>(c, 0);
      float* dv = corr_buf<float>(c);
      DGMI_CHECK(r32 && dv, "device allocation failed");
      // scaled residual in FP32; the correction is scaled back in FP64
      DGMI_TRY((kry::cast<double, float>(n, r64, r32, 1.0 / rnorm, s)));
      DGMI_TRY(gmres_cycle<float>(c, r32, dv, m, std::max(eps * 1.0, 1e-6), &used, s));
      DGMI_TRY(kry::add_cast<float>(n, dv, x, k->red + 5, s));  // red[5] = ||r||
    }
    its += used;
    ++outer;
    ++cycles;
  }
  if (stats) {
    stats[0] = its;
    stats[1] = bnorm > 0 ? rnorm / bnorm : 0.0;
    stats[2] = outer;
    stats[3] = cycles;
  }
  return 0;
}

int dgmi_set_param(dgmi_ctx* c, const char* name, double value) {
  DGMI_CHECK(c && name, "null argument");
  const std::string k(name);
  if (k == "gmres_orth") {
    DGMI_CHECK(value == 0.0 || value == 1.0, "gmres_orth must be 0 (MGS) or 1 (CGS2)");
    c->gmres_orth = (int)value;
    return 0;
  }
  set_error("unknown parameter " + k);
  return -2;
}

int dgmi_dot(int dtype, int64_t n, const void* x, const void* y, double* out, void* stream) {
  DGMI_CHECK(x && y && out, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64) return launch_dot<double>(n, (const double*)x, (const double*)y, out, 0, s);
  if (dtype == DGMI_F32) return launch_dot<float>(n, (const float*)x, (const float*)y, out, 0, s);
  set_error("dtype not supported");
  return -2;
}

int dgmi_axpy(int dtype, int64_t n, double a, const void* x, void* y, void* stream) {
  DGMI_CHECK(x && y, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == DGMI_F64) return launch_axpy<double>(n, a, (const double*)x, (double*)y, s);
  if (dtype == DGMI_F32) return launch_axpy<float>(n, a, (const float*)x, (float*)y, s);
  set_error("dtype not supported");
  return -2;
}

}  // extern "C"

// ========================================================================================
// Host-only set-up entry points (no device needed): C-ABI replacements of the reference's
// f2py Fortran modules and Python set-up loops.
extern "C" {

int dgmi_host_element_tables(int D, double* n, double* nlx, double* weight, double* sn,
                             double* snlx, double* sweight, int64_t* gi_align) {
  DGMI_CHECK(D == 2 || D == 3, "D must be 2 or 3");
  const ElementTables e = make_element(D);
  if (n) std::copy(e.n.begin(), e.n.end(), n);
  if (nlx) std::copy(e.nlx.begin(), e.nlx.end(), nlx);
  if (weight) std::copy(e.weight.begin(), e.weight.end(), weight);
  if (sn) std::copy(e.sn.begin(), e.sn.end(), sn);
  if (snlx) std::copy(e.snlx.begin(), e.snlx.end(), snlx);
  if (sweight) std::copy(e.sweight.begin(), e.sweight.end(), sweight);
  if (gi_align) std::copy(e.gi_align.begin(), e.gi_align.end(), gi_align);
  return 0;
}

int dgmi_host_topology(const int64_t* cells, int64_t nele, int D, int64_t* nbf, int64_t* alnmt,
                       int64_t* bctype) {
  DGMI_CHECK(cells && (D == 2 || D == 3), "bad argument");
  const Topology t = build_topology(cells, nele, D);
  std::copy(t.nbf.begin(), t.nbf.end(), nbf);
  std::copy(t.alnmt.begin(), t.alnmt.end(), alnmt);
  std::copy(t.bctype.begin(), t.bctype.end(), bctype);
  return 0;
}

int dgmi_host_p1cg_sparsity(const int64_t* cells, int64_t nele, int D, int64_t nverts,
                            int64_t cap, int64_t* fina, int64_t* cola, int64_t* ncola) {
  std::vector<int64_t> fa, ca;
  p1cg_sparsity(cells, nele, D, nverts, fa, ca);
  *ncola = (int64_t)ca.size();
  DGMI_CHECK((int64_t)ca.size() <= cap, "cola capacity too small");
  std::copy(fa.begin(), fa.end(), fina);
  std::copy(ca.begin(), ca.end(), cola);
  return 0;
}

int dgmi_host_color2(const int64_t* fina, const int64_t* cola, int64_t n, int32_t* whichc,
                     int32_t* ncolor) {
  std::vector<int64_t> fa(fina, fina + n + 1), ca(cola, cola + fina[n]);
  std::vector<int32_t> wc;
  *ncolor = color2(fa, ca, wc);
  std::copy(wc.begin(), wc.end(), whichc);
  return 0;
}

int dgmi_host_sfc_order(const double* verts, int64_t nverts, int D, int64_t* sfc) {
  DGMI_CHECK(verts && sfc && (D == 2 || D == 3), "bad argument");
  std::vector<int64_t> s;
  hilbert_order(verts, nverts, D, s);
  std::copy(s.begin(), s.end(), sfc);
  return 0;
}

int dgmi_host_nlevel_sfc(int64_t n) { return nlevel_sfc(n); }

int dgmi_host_sfc_levels(const int64_t* fina, const int64_t* cola, const double* a,
                         const int64_t* sfc, int64_t n, int64_t cap_rows, int64_t cap_nnz,
                         int64_t* nlevel, int64_t* level_n, int64_t* rowptr_all,
                         int64_t* col_all, double* val_all) {
  std::vector<int64_t> fa(fina, fina + n + 1), ca(cola, cola + fina[n]), sf(sfc, sfc + n);
  std::vector<double> av(a, a + fina[n]);
  SfcLevelsHost lv;
  build_sfc_levels(fa, ca, av, sf, lv);
  *nlevel = lv.nlevel;
  int64_t rows = 0, nnz = 0;
  for (int l = 0; l < lv.nlevel; ++l) {
    rows += lv.n[l] + 1;
    nnz += (int64_t)lv.col[l].size();
  }
  DGMI_CHECK(rows <= cap_rows && nnz <= cap_nnz, "level capacity too small");
  int64_t ro = 0, no = 0;
  for (int l = 0; l < lv.nlevel; ++l) {
    level_n[l] = lv.n[l];
    for (int64_t i = 0; i <= lv.n[l]; ++i) rowptr_all[ro + i] = lv.rowptr[l][i];
    for (size_t q = 0; q < lv.col[l].size(); ++q) {
      col_all[no + q] = lv.col[l][q];
      val_all[no + q] = lv.val[l][q];
    }
    ro += lv.n[l] + 1;
    no += (int64_t)lv.col[l].size();
  }
  return 0;
}

}  // extern "C"
