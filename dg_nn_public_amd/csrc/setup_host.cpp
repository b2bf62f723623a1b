// Host-side set-up: element tables, topology, geometry, P1CG graph, colouring, SFC.
#include "setup_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

namespace dgmi {
namespace {

constexpr int P = 3;

// barycentric multi-indices (sum P) in the reference's local node order
// (z02/shape_function.py:511-532 tets, :1127-1140 triangles)
const int TET_INDEX[20][4] = {
    {3, 0, 0, 0}, {0, 3, 0, 0}, {0, 0, 3, 0}, {0, 0, 0, 3}, {2, 0, 1, 0},
    {1, 0, 2, 0}, {2, 1, 0, 0}, {1, 2, 0, 0}, {0, 2, 1, 0}, {0, 1, 2, 0},
    {2, 0, 0, 1}, {1, 0, 0, 2}, {0, 2, 0, 1}, {0, 1, 0, 2}, {0, 0, 2, 1},
    {0, 0, 1, 2}, {0, 1, 1, 1}, {1, 1, 1, 0}, {1, 0, 1, 1}, {1, 1, 0, 1}};
const int TRI_INDEX[10][3] = {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0},
                              {0, 2, 1}, {0, 1, 2}, {1, 0, 2}, {2, 0, 1}, {1, 1, 1}};
// face vertex lists (z02/mesh_init.py:305-320, :43-45)
const int TET_FACE_VERTS[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
const int TRI_FACE_VERTS[3][2] = {{0, 1}, {1, 2}, {2, 0}};
// d(xi)/d(face params) (z02/function_space.py:200-213)
const double TET_DRST[4][3][2] = {{{0, 0}, {0, 1}, {1, 0}},
                                  {{1, 0}, {0, 0}, {0, 1}},
                                  {{0, 1}, {1, 0}, {0, 0}},
                                  {{-1, -1}, {1, 0}, {0, 1}}};
const double TRI_DRST[3][2] = {{-1.0, 1.0}, {0.0, -1.0}, {1.0, 0.0}};

// distinct permutations of a 4-tuple, in the order itertools.permutations yields them
void orbit4(const std::array<double, 4>& v, std::vector<std::array<double, 4>>& out) {
  std::array<int, 4> idx = {0, 1, 2, 3};
  std::vector<std::array<double, 4>> seen;
  do {
    std::array<double, 4> t = {v[idx[0]], v[idx[1]], v[idx[2]], v[idx[3]]};
    bool dup = false;
    for (auto& s : seen)
      if (s == t) { dup = true; break; }
    if (!dup) seen.push_back(t);
  } while (std::next_permutation(idx.begin(), idx.end()));
  out.insert(out.end(), seen.begin(), seen.end());
}

// Keast #7 (z02/shape_function.py:25-64): returns barycentric (ngi, 4) and weights
void tet_rule_24(std::vector<double>& lam, std::vector<double>& w) {
  struct Orb { std::array<double, 4> g; double w; };
  const Orb orbs[4] = {
      {{0.3561913862225449, 0.2146028712591517, 0.2146028712591517, 0.2146028712591517}, 0.0399227502581679},
      {{0.8779781243961660, 0.0406739585346113, 0.0406739585346113, 0.0406739585346113}, 0.0100772110553207},
      {{0.0329863295731731, 0.3223378901422757, 0.3223378901422757, 0.3223378901422757}, 0.0553571815436544},
      {{0.2696723314583159, 0.6030056647916491, 0.0636610018750175, 0.0636610018750175}, 0.0482142857142857}};
  lam.clear();
  w.clear();
  for (auto& o : orbs) {
    std::vector<std::array<double, 4>> pts;
    orbit4(o.g, pts);
    for (auto& p : pts) {
      double l1 = p[0], l2 = p[1], l3 = p[2];
      lam.push_back(l1);
      lam.push_back(l2);
      lam.push_back(l3);
      lam.push_back(1.0 - (l1 + l2 + l3));
      w.push_back(o.w / 6.0);
    }
  }
}

// Strang-Fix #9 (z02/shape_function.py:271-308): barycentric (12, 3), weights sum 1/2
void tri_rule_12(std::vector<double>& pts, std::vector<double>& w) {
  const double a = 0.873821971016996, b = 0.063089014491502;
  const double c = 0.501426509658179, d = 0.249286745170910;
  const double e = 0.636502499121399, f = 0.310352451033785, g = 0.053145049844816;
  std::vector<std::array<double, 2>> xy = {{a, b}, {b, a}, {b, b}, {c, d}, {d, c}, {d, d}};
  std::array<double, 3> efg = {e, f, g};
  std::array<int, 3> idx = {0, 1, 2};
  do {
    xy.push_back({efg[idx[0]], efg[idx[1]]});
  } while (std::next_permutation(idx.begin(), idx.end()));
  pts.clear();
  for (auto& p : xy) {
    pts.push_back(p[0]);
    pts.push_back(p[1]);
    pts.push_back(1.0 - (p[0] + p[1]));
  }
  w.assign(12, 0.0);
  for (int i = 0; i < 3; ++i) w[i] = 0.050844906370207 * 0.5;
  for (int i = 3; i < 6; ++i) w[i] = 0.116786275726379 * 0.5;
  for (int i = 6; i < 12; ++i) w[i] = 0.082851075618374 * 0.5;
}

// 4-point Gauss-Legendre on [0,1] (z02/shape_function.py:392-420)
void line_rule_4(std::vector<double>& pts, std::vector<double>& w) {
  const double a = 0.339981043584856, b = 0.861136311594053;
  const double w1 = 0.652145154862546, w2 = 0.347854845137454;
  const double s[4] = {0.5 - 0.5 * b, 0.5 - 0.5 * a, 0.5 + 0.5 * a, 0.5 + 0.5 * b};
  pts.clear();
  for (int i = 0; i < 4; ++i) {
    pts.push_back(s[i]);
    pts.push_back(1.0 - s[i]);
  }
  w = {w2 / 2.0, w1 / 2.0, w1 / 2.0, w2 / 2.0};
}

// f_k(t) = prod_{j<k} (P t - j)/(j+1) and derivative
inline void factor(int k, double t, double& val, double& der) {
  val = 1.0;
  der = 0.0;
  for (int j = 0; j < k; ++j) {
    double term = (P * t - j) / (j + 1);
    der = der * term + val * (double(P) / (j + 1));
    val = val * term;
  }
}

// phi (nloc) and dphi/dxi (D, nloc) at one barycentric point
void lagrange(const std::vector<int>& index, int nb, const double* lam, double* phi,
              double* dxi, int nloc) {
  int D = nb - 1;
  std::vector<double> vals(nb), ders(nb), dl(nb);
  for (int n = 0; n < nloc; ++n) {
    for (int i = 0; i < nb; ++i) factor(index[n * nb + i], lam[i], vals[i], ders[i]);
    double p = 1.0;
    for (int i = 0; i < nb; ++i) p *= vals[i];
    phi[n] = p;
    for (int i = 0; i < nb; ++i) {
      double prod = ders[i];
      for (int k = 0; k < nb; ++k)
        if (k != i) prod *= vals[k];
      dl[i] = prod;
    }
    for (int d = 0; d < D; ++d) dxi[d * nloc + n] = dl[d] - dl[D];
  }
}

// element barycentrics of the point with face barycentrics p on face f
// (z02/shape_function.py:763-783 tets, :1105-1116 triangles)
void face_lam(int D, int f, const double* p, double* lam) {
  for (int i = 0; i < 4; ++i) lam[i] = 0.0;
  if (D == 3) {
    const double p0 = p[0], p1 = p[1], p2 = p[2];
    const double tab[4][4] = {{0, p1, p0, p2}, {p0, 0, p1, p2}, {p1, p0, 0, p2}, {p0, p1, p2, 0}};
    for (int i = 0; i < 4; ++i) lam[i] = tab[f][i];
  } else {
    const double s0 = p[0], s1 = p[1];
    const double tab[3][3] = {{s0, s1, 0}, {0, s0, s1}, {s1, 0, s0}};
    for (int i = 0; i < 3; ++i) lam[i] = tab[f][i];
  }
}

}  // namespace

ElementTables make_element(int D) {
  ElementTables el;
  el.D = D;
  el.nface = D + 1;
  const int nb = D + 1;
  std::vector<double> vlam, vw, fp, fw;
  if (D == 3) {
    el.nloc = 20;
    el.snloc = 10;
    for (int n = 0; n < 20; ++n)
      for (int i = 0; i < 4; ++i) el.index.push_back(TET_INDEX[n][i]);
    tet_rule_24(vlam, vw);
    tri_rule_12(fp, fw);
  } else {
    el.nloc = 10;
    el.snloc = 4;
    for (int n = 0; n < 10; ++n)
      for (int i = 0; i < 3; ++i) el.index.push_back(TRI_INDEX[n][i]);
    tri_rule_12(vlam, vw);
    line_rule_4(fp, fw);
  }
  const int nloc = el.nloc;
  el.ngi = (int)vw.size();
  el.sngi = (int)fw.size();
  el.weight = vw;
  el.sweight = fw;
  el.n.assign(nloc * el.ngi, 0.0);
  el.nlx.assign(D * nloc * el.ngi, 0.0);
  std::vector<double> phi(nloc), dxi(D * nloc);
  for (int g = 0; g < el.ngi; ++g) {
    lagrange(el.index, nb, &vlam[g * nb], phi.data(), dxi.data(), nloc);
    for (int n = 0; n < nloc; ++n) {
      el.n[n * el.ngi + g] = phi[n];
      for (int d = 0; d < D; ++d) el.nlx[(d * nloc + n) * el.ngi + g] = dxi[d * nloc + n];
    }
  }
  // face points in element barycentrics
  const int nf = el.nface, sg = el.sngi;
  el.sn.assign(nf * nloc * sg, 0.0);
  el.snlx.assign(nf * D * nloc * sg, 0.0);
  for (int f = 0; f < nf; ++f) {
    for (int g = 0; g < sg; ++g) {
      double lam[4];
      face_lam(D, f, &fp[g * D], lam);
      lagrange(el.index, nb, lam, phi.data(), dxi.data(), nloc);
      for (int n = 0; n < nloc; ++n) {
        el.sn[(f * nloc + n) * sg + g] = phi[n];
        for (int d = 0; d < D; ++d)
          el.snlx[((f * D + d) * nloc + n) * sg + g] = dxi[d * nloc + n];
      }
    }
  }
  // alignment: face point g on this side has weights fp[g] on the face's vertex list;
  // the neighbour lists the same vertices reversed and rotated to start at this side's
  // vertex k (alnmt = k).  Row k maps g to the neighbour's matching point.
  el.gi_align.assign(D * sg, 0);
  for (int k = 0; k < D; ++k) {
    int nb_order[3];
    if (D == 3) {
      const int rev[3] = {0, 2, 1};
      int start = 0;
      for (int i = 0; i < 3; ++i)
        if (rev[i] == k) start = i;
      for (int i = 0; i < 3; ++i) nb_order[i] = rev[(start + i) % 3];
    } else {
      nb_order[0] = (k == 0) ? 0 : 1;
      nb_order[1] = (k == 0) ? 1 : 0;
    }
    for (int g = 0; g < sg; ++g) {
      int best = -1;
      double bestd = 1e300;
      for (int h = 0; h < sg; ++h) {
        double d = 0.0;
        for (int i = 0; i < D; ++i) {
          // neighbour point h weight on this-side vertex nb_order[i] is fp[h][i]
          double wn = fp[h * D + i];
          double wt = fp[g * D + nb_order[i]];
          d += std::fabs(wn - wt);
        }
        if (d < bestd) { bestd = d; best = h; }
      }
      el.gi_align[k * sg + g] = best;
    }
  }
  // quadratic interpolation of face gradients: nodes in face barycentrics, their Lagrange
  // basis at the face points, gradients there; checked against the tabulated snlx
  {
    const int fb = D;  // face barycentric coordinates per point
    std::vector<std::array<double, 3>> qn;
    if (D == 3)
      qn = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0.5, 0.5, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}};
    else
      qn = {{1, 0, 0}, {0, 1, 0}, {0.5, 0.5, 0}};
    const int nk = (int)qn.size();
    el.nfk = nk;
    auto quad_basis = [&](const double* mu, double* L) {
      if (D == 3) {
        for (int i = 0; i < 3; ++i) L[i] = mu[i] * (2.0 * mu[i] - 1.0);
        L[3] = 4.0 * mu[0] * mu[1];
        L[4] = 4.0 * mu[1] * mu[2];
        L[5] = 4.0 * mu[0] * mu[2];
      } else {
        L[0] = mu[0] * (2.0 * mu[0] - 1.0);
        L[1] = mu[1] * (2.0 * mu[1] - 1.0);
        L[2] = 4.0 * mu[0] * mu[1];
      }
    };
    el.flk.assign(sg * nk, 0.0);
    for (int g = 0; g < sg; ++g) quad_basis(&fp[g * fb], &el.flk[g * nk]);
    el.fgk.assign(nf * nk * D * nloc, 0.0);
    el.fnodes.clear();
    el.fphi.assign(nf * sg * el.snloc, 0.0);
    double err = 0.0, big = 0.0;
    for (int f = 0; f < nf; ++f) {
      for (int k = 0; k < nk; ++k) {
        double lam[4];
        face_lam(D, f, qn[k].data(), lam);
        lagrange(el.index, nb, lam, phi.data(), dxi.data(), nloc);
        for (int d = 0; d < D; ++d)
          for (int n = 0; n < nloc; ++n) el.fgk[((f * nk + k) * D + d) * nloc + n] = dxi[d * nloc + n];
      }
      std::vector<int> on;
      for (int n = 0; n < nloc; ++n) {
        double m = 0.0;
        for (int g = 0; g < sg; ++g) m = std::max(m, std::fabs(el.sn[(f * nloc + n) * sg + g]));
        if (m != 0.0) on.push_back(n);
      }
      if ((int)on.size() != el.snloc) err = INFINITY;
      for (int i = 0; i < (int)on.size() && i < el.snloc; ++i) {
        el.fnodes.push_back(on[i]);
        for (int g = 0; g < sg; ++g)
          el.fphi[(f * sg + g) * el.snloc + i] = el.sn[(f * nloc + on[i]) * sg + g];
      }
      for (int d = 0; d < D; ++d)
        for (int n = 0; n < nloc; ++n)
          for (int g = 0; g < sg; ++g) {
            double v = 0.0;
            for (int k = 0; k < nk; ++k)
              v += el.flk[g * nk + k] * el.fgk[((f * nk + k) * D + d) * nloc + n];
            const double ref = el.snlx[((f * D + d) * nloc + n) * sg + g];
            err = std::max(err, std::fabs(v - ref));
            big = std::max(big, std::fabs(ref));
          }
    }
    el.face_interp_err = big > 0.0 ? err / big : err;
  }
  el.p1_prolong.assign(nloc * nb, 0.0);
  for (int n = 0; n < nloc; ++n)
    for (int k = 0; k < nb; ++k) el.p1_prolong[n * nb + k] = el.index[n * nb + k] / double(P);
  return el;
}

// elements of every vertex, in increasing element order (CSR)
static void vertex_elements(const int64_t* cells, int64_t nele, int nb, int64_t nverts,
                            std::vector<int64_t>& vptr, std::vector<int64_t>& velem) {
  vptr.assign(nverts + 1, 0);
  for (int64_t q = 0; q < nele * nb; ++q) vptr[cells[q] + 1]++;
  for (int64_t v = 0; v < nverts; ++v) vptr[v + 1] += vptr[v];
  velem.resize(nele * nb);
  std::vector<int64_t> fill(vptr.begin(), vptr.end() - 1);
  for (int64_t e = 0; e < nele; ++e)
    for (int k = 0; k < nb; ++k) velem[fill[cells[e * nb + k]]++] = e;
}

Topology build_topology(const int64_t* cells, int64_t nele, int D) {
  // The neighbour across face (e, f) is the other element holding the face's vertices:
  // scan the elements of the face's smallest vertex.  The alignment of a matched pair is
  // read from the pair (lower global face id first) and stored on both sides.
  const int nf = D + 1, nb = D + 1;
  const int64_t nfaces = nele * nf;
  Topology t;
  t.nbf.assign(nfaces, -1);
  t.alnmt.assign(nfaces, -1);
  t.bctype.assign(nfaces, 0);
  int64_t nverts = 0;
  for (int64_t q = 0; q < nele * nb; ++q) nverts = std::max(nverts, cells[q] + 1);
  std::vector<int64_t> vptr, velem;
  vertex_elements(cells, nele, nb, nverts, vptr, velem);
  auto face_verts = [&](int64_t e, int f, int64_t* v) {
    for (int i = 0; i < D; ++i)
      v[i] = cells[e * nb + (D == 3 ? TET_FACE_VERTS[f][i] : TRI_FACE_VERTS[f][i])];
  };
  parallel_chunks(nele, [&](int, int64_t e0, int64_t e1) {
    for (int64_t e = e0; e < e1; ++e)
      for (int f = 0; f < nf; ++f) {
        int64_t v[3] = {0, 0, 0}, s3[3] = {0, 0, 0};
        face_verts(e, f, v);
        std::copy(v, v + D, s3);
        std::sort(s3, s3 + D);
        for (int64_t q = vptr[s3[0]]; q < vptr[s3[0] + 1]; ++q) {
          const int64_t e2 = velem[q];
          if (e2 == e) continue;
          int hit = 0;
          for (int k = 0; k < nb; ++k)
            for (int i = 1; i < D; ++i) hit += cells[e2 * nb + k] == s3[i];
          if (hit != D - 1) continue;
          for (int f2 = 0; f2 < nf; ++f2) {
            int64_t w[3] = {0, 0, 0}, ws[3] = {0, 0, 0};
            face_verts(e2, f2, w);
            std::copy(w, w + D, ws);
            std::sort(ws, ws + D);
            if (!std::equal(ws, ws + D, s3)) continue;
            const int64_t z = e * nf + f, z2 = e2 * nf + f2;
            const int64_t* va = z < z2 ? v : w;
            const int64_t* vb = z < z2 ? w : v;
            t.nbf[z] = z2;
            t.alnmt[z] = va[0] == vb[0] ? 0 : (va[1] == vb[0] ? 1 : 2);
            break;
          }
          break;
        }
      }
  });
  for (int64_t z = 0; z < nfaces; ++z) t.bctype[z] = t.nbf[z] >= 0 ? -1 : 0;
  return t;
}

bool build_geometry(const ElementTables& el, const double* x, const double* kappa,
                    const Topology& topo, int64_t nele, double eta_e, int face_kappa, Geometry& g,
                    std::string& err) {
  const int D = el.D, nloc = el.nloc, nf = el.nface, nb = D + 1;
  const int ng = D == 3 ? 6 : 3;
  g.G.assign(ng * nele, 0.0);
  g.cf.assign(nf * D * nele, 0.0);
  g.wf.assign(nf * nele, 0.0);
  g.pen.assign(nf * nele, 0.0);
  g.nbface.assign(nf * nele, -1);
  g.fcode.assign(nf * nele, 0);
  g.detJ.assign(nele, 0.0);
  g.max_affine_dev = 0.0;
  double swsum = 0.0;
  for (double w : el.sweight) swsum += w;
  // elements are independent: chunks per host thread, per-thread deviation and error
  const int nth = host_threads();
  std::vector<double> tdev(nth, 0.0);
  std::vector<int64_t> tbad(nth, -1);
  parallel_chunks(nele, [&](int tid, int64_t e0, int64_t e1) {
  for (int64_t e = e0; e < e1; ++e) {
    const double* xe = x + e * D * nloc;  // (D, nloc)
    double v[4][3] = {{0}};
    for (int k = 0; k < nb; ++k)
      for (int j = 0; j < D; ++j) v[k][j] = xe[j * nloc + k];
    double J[3][3] = {{0}};
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) J[i][j] = v[i][j] - v[D][j];
    // affinity check of the P3 nodes
    double hscale = 0.0;
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) hscale = std::max(hscale, std::fabs(J[i][j]));
    for (int n = 0; n < nloc; ++n)
      for (int j = 0; j < D; ++j) {
        double exp_x = 0.0;
        for (int k = 0; k < nb; ++k) exp_x += el.p1_prolong[n * nb + k] * v[k][j];
        double dev = std::fabs(xe[j * nloc + n] - exp_x) / (hscale > 0 ? hscale : 1.0);
        tdev[tid] = std::max(tdev[tid], dev);
      }
    double det, inv[3][3];
    if (D == 3) {
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
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
      if (tbad[tid] < 0) tbad[tid] = e;
      continue;
    }
    g.detJ[e] = det;
    const double ke = kappa[e];
    // metric kappa detJ invJ^T invJ  (G[j][k] = sum_i inv[i][j] inv[i][k])
    auto Gjk = [&](int j, int k) {
      double s = 0.0;
      for (int i = 0; i < D; ++i) s += inv[i][j] * inv[i][k];
      return ke * det * s;
    };
    if (D == 3) {
      const double gv[6] = {Gjk(0, 0), Gjk(1, 1), Gjk(2, 2), Gjk(0, 1), Gjk(0, 2), Gjk(1, 2)};
      for (int c = 0; c < 6; ++c) g.G[c * nele + e] = gv[c];
    } else {
      const double gv[3] = {Gjk(0, 0), Gjk(1, 1), Gjk(0, 1)};
      for (int c = 0; c < 3; ++c) g.G[c * nele + e] = gv[c];
    }
    for (int f = 0; f < nf; ++f) {
      double nrm[3] = {0, 0, 0}, sdet;
      if (D == 3) {
        double t[2][3];
        for (int k = 0; k < 2; ++k)
          for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i) s += J[i][j] * TET_DRST[f][i][k];
            t[k][j] = s;
          }
        double cr[3] = {t[0][1] * t[1][2] - t[0][2] * t[1][1],
                        t[0][2] * t[1][0] - t[0][0] * t[1][2],
                        t[0][0] * t[1][1] - t[0][1] * t[1][0]};
        sdet = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int j = 0; j < 3; ++j) nrm[j] = cr[j] / sdet;
      } else {
        double t[2];
        for (int j = 0; j < 2; ++j) {
          double s = 0.0;
          for (int i = 0; i < 2; ++i) s += J[i][j] * TRI_DRST[f][i];
          t[j] = s;
        }
        sdet = std::sqrt(t[0] * t[0] + t[1] * t[1]);
        nrm[0] = t[1] / sdet;
        nrm[1] = -t[0] / sdet;
      }
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
        for (int i = 0; i < D; ++i) s += inv[i][j] * nrm[i];
        g.cf[(f * D + j) * nele + e] = s;
      }
      double h = sdet * swsum;
      if (D == 3) h = std::sqrt(h);
      g.pen[f * nele + e] = eta_e / h;
      const int64_t z = e * nf + f;
      const int64_t type = topo.bctype[z];
      double kf = 0.0;
      if (type < 0 && topo.nbf[z] >= 0) {
        const int64_t nbe = topo.nbf[z] / nf;
        const double kn = kappa[nbe];
        kf = face_kappa == 1 ? 2.0 * ke * kn / (ke + kn) : 0.5 * (ke + kn);
        g.nbface[z] = (int32_t)topo.nbf[z];
        g.fcode[z] = 0 | ((int32_t)topo.alnmt[z] << 4);
      } else if (type == 0) {
        kf = ke;
        g.fcode[z] = 1;
      } else if (type == 1) {
        // Neumann: no operator term; the right-hand side adds int v g_N without kappa
        // (z02/volume_mf_diff.py:360-400)
        kf = 1.0;
        g.fcode[z] = 3;
      } else {
        kf = 0.0;
        g.fcode[z] = 2;
      }
      g.wf[f * nele + e] = kf * sdet;
    }
  }
  });
  for (int t = 0; t < nth; ++t) {
    g.max_affine_dev = std::max(g.max_affine_dev, tdev[t]);
    if (tbad[t] >= 0) {
      err = "degenerate element " + std::to_string(tbad[t]);
      return false;
    }
  }
  return true;
}

void p1cg_sparsity(const int64_t* cells, int64_t nele, int D, int64_t nverts,
                   std::vector<int64_t>& fina, std::vector<int64_t>& cola) {
  // row i = sorted distinct vertices of the elements around vertex i
  const int nb = D + 1;
  std::vector<int64_t> vptr, velem;
  vertex_elements(cells, nele, nb, nverts, vptr, velem);
  auto row = [&](int64_t v, std::vector<int64_t>& buf) {
    buf.clear();
    for (int64_t q = vptr[v]; q < vptr[v + 1]; ++q)
      for (int k = 0; k < nb; ++k) buf.push_back(cells[velem[q] * nb + k]);
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  fina.assign(nverts + 1, 0);
  parallel_chunks(nverts, [&](int, int64_t a, int64_t b) {
    std::vector<int64_t> buf;
    for (int64_t v = a; v < b; ++v) {
      row(v, buf);
      fina[v + 1] = (int64_t)buf.size();
    }
  });
  for (int64_t v = 0; v < nverts; ++v) fina[v + 1] += fina[v];
  cola.resize(fina[nverts]);
  parallel_chunks(nverts, [&](int, int64_t a, int64_t b) {
    std::vector<int64_t> buf;
    for (int64_t v = a; v < b; ++v) {
      row(v, buf);
      std::copy(buf.begin(), buf.end(), cola.begin() + fina[v]);
    }
  });
}

int color2(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
           std::vector<int32_t>& whichc) {
  const int64_t n = (int64_t)fina.size() - 1;
  whichc.assign(n, 0);
  int ncolor = 0;
  std::vector<int64_t> todo(n);
  std::iota(todo.begin(), todo.end(), 0);
  while (!todo.empty()) {
    ++ncolor;
    for (int64_t i : todo) {
      bool clash = false;
      for (int64_t c = fina[i]; c < fina[i + 1] && !clash; ++c) {
        int64_t j = cola[c];
        if (whichc[j] == ncolor) { clash = true; break; }
        for (int64_t c2 = fina[j]; c2 < fina[j + 1]; ++c2)
          if (whichc[cola[c2]] == ncolor) { clash = true; break; }
      }
      if (!clash) whichc[i] = ncolor;
    }
    std::vector<int64_t> next;
    for (int64_t i : todo)
      if (whichc[i] == 0) next.push_back(i);
    todo.swap(next);
  }
  return ncolor;
}

void hilbert_order(const double* verts, int64_t nv, int D, std::vector<int64_t>& sfc) {
  const int bits = D == 3 ? 10 : 16;
  double lo[3], hi[3];
  for (int d = 0; d < D; ++d) {
    lo[d] = 1e300;
    hi[d] = -1e300;
  }
  for (int64_t i = 0; i < nv; ++i)
    for (int d = 0; d < D; ++d) {
      lo[d] = std::min(lo[d], verts[i * D + d]);
      hi[d] = std::max(hi[d], verts[i * D + d]);
    }
  const double scale = double((1LL << bits) - 1);
  std::vector<int64_t> key(nv);
  for (int64_t i = 0; i < nv; ++i) {
    int64_t X[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
      double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
      double t = (verts[i * D + d] - lo[d]) / span;
      t = t * scale;
      t = t + 0.5;
      X[d] = (int64_t)std::floor(t);
    }
    // Skilling, "Programming the Hilbert curve" (AIP Conf. Proc. 707, 2004)
    const int64_t M = 1LL << (bits - 1);
    for (int64_t Q = M; Q > 1; Q >>= 1) {
      int64_t Pm = Q - 1;
      for (int k = 0; k < D; ++k) {
        if (X[k] & Q) {
          X[0] ^= Pm;
        } else {
          int64_t t = (X[0] ^ X[k]) & Pm;
          X[0] ^= t;
          X[k] ^= t;
        }
      }
    }
    for (int k = 1; k < D; ++k) X[k] ^= X[k - 1];
    int64_t t = 0;
    for (int64_t Q = M; Q > 1; Q >>= 1)
      if (X[D - 1] & Q) t ^= Q - 1;
    for (int k = 0; k < D; ++k) X[k] ^= t;
    int64_t kk = 0;
    for (int j = bits - 1; j >= 0; --j)
      for (int k = 0; k < D; ++k) kk = (kk << 1) | ((X[k] >> j) & 1);
    key[i] = kk;
  }
  std::vector<int64_t> order(nv);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  sfc.assign(nv, 0);
  for (int64_t p = 0; p < nv; ++p) sfc[order[p]] = p + 1;
}

int nlevel_sfc(int64_t n) {
  if (n == 1) return 1;
  int lev = 1;
  while (n > 1) {
    n = (n - 1) / 2 + 1;
    ++lev;
  }
  return lev;
}

void build_sfc_levels(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
                      const std::vector<double>& rar, const std::vector<int64_t>& sfc,
                      SfcLevelsHost& lv) {
  const int64_t n0 = (int64_t)fina.size() - 1;
  lv.nlevel = nlevel_sfc(n0);
  lv.n.clear();
  lv.rowptr.clear();
  lv.col.clear();
  lv.val.clear();
  lv.diag.clear();
  // level 0 in SFC order: entries (pos(i), pos(j), a_ij)
  struct Ent { int64_t r, c; double v; };
  std::vector<Ent> ents;
  ents.reserve(cola.size());
  for (int64_t i = 0; i < n0; ++i)
    for (int64_t c = fina[i]; c < fina[i + 1]; ++c)
      ents.push_back({sfc[i] - 1, sfc[cola[c]] - 1, rar[c]});
  int64_t nl = n0;
  for (int l = 0; l < lv.nlevel; ++l) {
    if (l > 0) {
      for (auto& e : ents) {
        e.r >>= 1;
        e.c >>= 1;
      }
      nl = (nl - 1) / 2 + 1;
    }
    std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
      return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    std::vector<Ent> red;
    red.reserve(ents.size());
    for (auto& e : ents) {
      if (!red.empty() && red.back().r == e.r && red.back().c == e.c) red.back().v += e.v;
      else red.push_back(e);
    }
    ents.swap(red);
    std::vector<int32_t> rp(nl + 1, 0), cl(ents.size());
    std::vector<double> vv(ents.size()), dg(nl, 0.0);
    for (size_t k = 0; k < ents.size(); ++k) {
      rp[ents[k].r + 1]++;
      cl[k] = (int32_t)ents[k].c;
      vv[k] = ents[k].v;
      if (ents[k].r == ents[k].c) dg[ents[k].r] = ents[k].v;
    }
    for (int64_t i = 0; i < nl; ++i) rp[i + 1] += rp[i];
    lv.n.push_back(nl);
    lv.rowptr.push_back(std::move(rp));
    lv.col.push_back(std::move(cl));
    lv.val.push_back(std::move(vv));
    lv.diag.push_back(std::move(dg));
  }
}

}  // namespace dgmi
