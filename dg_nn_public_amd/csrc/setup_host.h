// Host-side set-up for the DG diffusion context: reference-element tables, mesh
// topology, affine element geometry, P1CG graph, SFC ordering and SFC level operators.
// Pure C++ (no HIP); compiled with -ffp-contract=off so that the SFC quantisation and
// the tables evaluate in the same IEEE operation order as the Python oracle.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

namespace dgmi {

// Host threads used by the set-up (DGMI_HOST_THREADS overrides; at most 16).
inline int host_threads() {
  static const int n = [] {
    if (const char* ev = std::getenv("DGMI_HOST_THREADS")) return std::max(1, std::atoi(ev));
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc));
  }();
  return n;
}

// fn(t, begin, end) over contiguous chunks of [0, n), one per thread; small ranges run
// inline.
template <typename F>
void parallel_chunks(int64_t n, F fn, int64_t min_chunk = 4096) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / min_chunk));
  if (nt <= 1) {
    fn(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t) {
    const int64_t a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([=, &fn] { fn(t, a, b); });
  }
  for (auto& x : th) x.join();
}

// Reference P3 element (the reference's SHATRInew outputs for p = 3, quadrature
// degree 6: z02/shape_function.py:459-1317, sizes z02/function_space.py:32-66).
struct ElementTables {
  int D = 0, nloc = 0, ngi = 0, sngi = 0, nface = 0, snloc = 0;
  std::vector<double> n;        // (nloc, ngi)
  std::vector<double> nlx;      // (D, nloc, ngi)
  std::vector<double> weight;   // (ngi)
  std::vector<double> sn;       // (nface, nloc, sngi)
  std::vector<double> snlx;     // (nface, D, nloc, sngi)
  std::vector<double> sweight;  // (sngi)
  std::vector<int> gi_align;    // (D, sngi)
  std::vector<double> p1_prolong;  // (nloc, D+1)
  std::vector<int> face_verts;     // (nface, D) local vertices of each face
  // d(xi)/d(face params) per face (z02/function_space.py:200-213): 3D (nface, 3, 2),
  // 2D (nface, 2)
  std::vector<double> drst;
  std::vector<int> index;       // (nloc, D+1) barycentric multi-indices
  // Face traces through a quadratic interpolant.  On face f only the snloc nodes fnodes[f]
  // have nonzero basis functions, and the reference gradient of every basis function
  // restricted to the face is a quadratic polynomial, fixed by its values at the nfk
  // quadratic nodes of the face (vertices and edge midpoints: 6 in 3D, 3 in 2D).
  int nfk = 0;
  std::vector<int> fnodes;      // (nface, snloc) nodes on each face, ascending
  std::vector<double> fphi;     // (nface, sngi, snloc) phi of those nodes at the face points
  std::vector<double> flk;      // (sngi, nfk) quadratic face Lagrange basis at the face points
  std::vector<double> fgk;      // (nface, nfk, D, nloc) dphi/dxi at the face's quadratic nodes
  double face_interp_err = 0.0; // max |interpolated - tabulated| over snlx (set-up check)
  // Face terms in node form: with u on face f given by its face-node values and du/dn by
  // its values at the quadratic nodes, the lifts sum_g sw_g phi_i(g) (.)(g) become
  //   fmf (nface, snloc, snloc)  sum_g sw_g phi_i(g) phi_j(g)   (face mass matrix)
  //   fml (nface, snloc, nfk)    sum_g sw_g phi_i(g) L_k(g)
  std::vector<double> fmf, fml;
  // Volume term through the quadratic gradient: d phi_n / d xi_d is a P2 polynomial, fixed
  // by its values at the np2 P2 Lagrange nodes (vertices and edge midpoints), so
  //   sum_g w_g dphi_n/dxi_d(g) dphi_m/dxi_e(g) = sum_ij vd[d][i][n] vm2[i][j] vd[e][j][m]
  //   vd  (D, np2, nloc)  d phi_n / d xi_d at P2 node i
  //   vm2 (np2, np2)      sum_g w_g psi_i(g) psi_j(g)  (P2 mass matrix, exact for the rule)
  int np2 = 0;
  std::vector<double> vd, vm2;
  double vol_factor_err = 0.0;  // max |factored - tabulated| of those products (set-up check)
};
ElementTables make_element(int D);

// nbf / alnmt / bctype as z02/mesh_init.py:293-390 (3D) and :40-100 (2D).
struct Topology {
  std::vector<int64_t> nbf, alnmt, bctype;  // (nele*nface)
};
Topology build_topology(const int64_t* cells, int64_t nele, int D);

// P1CG CSR graph (z02/sparsity.py:102-122): all vertex pairs per element, sorted.
void p1cg_sparsity(const int64_t* cells, int64_t nele, int D, int64_t nverts,
                   std::vector<int64_t>& fina, std::vector<int64_t>& cola);
// greedy distance-2 colouring, colours 1..ncolor (z02/sparsity.py:235-269)
int color2(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
           std::vector<int32_t>& whichc);
// the same on the vertex range [v0, v1) of an independent block (whichc zero there before)
int color2_range(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola, int64_t v0,
                 int64_t v1, std::vector<int32_t>& whichc);
// Hilbert SFC numbering: sfc[node] = 1-based position (ties broken by node id).
void hilbert_order(const double* verts, int64_t nverts, int D, std::vector<int64_t>& sfc);
int nlevel_sfc(int64_t n);

// SFC level operators in CSR (SFC order), level l aggregates positions p >> l
// (z00-fortran-src/compact_fcode.f90:7-225).
struct SfcLevelsHost {
  int nlevel = 0;
  std::vector<int64_t> n;                 // nodes per level
  std::vector<std::vector<int32_t>> rowptr, col;
  std::vector<std::vector<double>> val, diag;
};
void build_sfc_levels(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
                      const std::vector<double>& rar, const std::vector<int64_t>& sfc,
                      SfcLevelsHost& lv);
// The same in two steps: the level patterns need only the P1CG pattern and the SFC order
// (so they can be formed before RAR is known), and record which entries each entry sums;
// filling the values from RAR then repeats build_sfc_levels' sums in its order.
struct SfcLevelPattern {
  int nlevel = 0;
  std::vector<int64_t> n;
  std::vector<std::vector<int32_t>> rowptr, col;
  std::vector<int32_t> perm0;  // level-0 entry k takes the input CSR entry perm0[k]
  // level l >= 1: entry k sums the level l-1 entries src[l][sptr[l][k] .. sptr[l][k+1])
  std::vector<std::vector<int32_t>> sptr, src;
};
void build_sfc_level_pattern(const std::vector<int64_t>& fina, const std::vector<int64_t>& cola,
                             const std::vector<int64_t>& sfc, SfcLevelPattern& pat);
// moves the pattern's rowptr and col into lv
void fill_sfc_levels(SfcLevelPattern& pat, const std::vector<double>& rar, SfcLevelsHost& lv);

}  // namespace dgmi
