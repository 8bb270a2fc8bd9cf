// Host-side setup for the adaptive sparse-grid DG hot path: 1D hierarchical
// bases, Gauss-Legendre products, level-ordered 1D operator tables, Lagrange
// point tables, hierarchical-surplus weights and the element hash key.
//
// Everything here runs once per problem (or once per grid change for the
// hash keys) on the host; the results are uploaded to the device as dense
// tables.  Formulas and their floating-point evaluation order follow the
// reference so that the tables are bit-identical to what it builds:
//   AlptBasis   source/AlptBasis.cpp:16-130,256-340,473-560
//   LagrBasis   source/LagrBasis.cpp:9-160,190-420,852-864
//   Basis       source/Basis.cpp:8-81,121-173 (product_function :38-59)
//   Quad::GL_1D source/Quad.cpp:31-80,224-244,140-146
//   AllBasis    include/AllBasis.h:53-78
//   OperatorMatrix1D include/OperatorMatrix1D.h:124-200
//   Lag_pt_Alpt_1D   source/Interplation.cpp:16-99
//   pw1d weights     source/Interplation.cpp:775-887
//   Hash        source/Hash.cpp:55-167
#pragma once
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

namespace adg {

constexpr double ROUND_OFF = 1e-13;  // include/libs.h:28
constexpr double PI = 3.1415926535897932384626433832795028841971693993751058209749445923078164062;

enum class BasisKind : int { Alpert = 0, Lagrange = 1, Hermite = 2 };

// 1D hierarchical index helpers.  "slot" is the level-ordered position of a
// 1D cell: level 0 -> 0, level n>=1 support j (odd) -> 2^(n-1) + (j-1)/2
// (Element::index_to_order_elem, source/Element.cpp:388-391).
inline int slot_of(int level, int suppt) { return level == 0 ? 0 : (1 << (level - 1)) + (suppt - 1) / 2; }
inline int level_of_slot(int s) { int n = 0; while ((1 << n) <= s) ++n; return n; }  // 0->0, 1->1, 2..3->2
inline int suppt_of_slot(int s) { if (s == 0) return 1; int n = level_of_slot(s); return 2 * (s - (1 << (n - 1))) + 1; }

// Support interval of the 1D cell (Basis::Basis, source/Basis.cpp:21-36).
void cell_support(int level, int suppt, double& xl, double& xr);

// A family of 1D basis functions up to level NMAX (AllBasis<T>).
struct Basis1D {
  BasisKind kind;
  int pmax;       // polynomial degree
  int msh_case;   // Lagrange point family (only msh_case 1 is provided)
  int nmax;       // maximum mesh level
  std::vector<double> msh0, msh1;  // Lagrange points on levels 0 and 1, mapped to [0,1]

  Basis1D(BasisKind k, int pmax, int nmax, int msh_case = 1);
  int size() const { return (1 << nmax) * (pmax + 1); }
  void decode(int order, int& level, int& suppt, int& p) const;
  // value (deriv 0/1) or one-sided limit (sgn -1/+1) of basis function `order` at x
  double val(int order, double x, int deriv = 0, int sgn = 0) const;
  // interpolation point of Lagrange basis `order` (LagrBasis::intep_pt)
  double point(int order) const;
  // Hermite degrees of freedom: order's p -> (point 0/1, derivative 0/1)
  // (HermBasis::deg_pt_deri_1d): 0 value at left, 1 value at right, 2 slope
  // at left, 3 slope at right
  static int herm_deriv(int p) { return p / 2; }
};

// GL 10-point rule on [tl,tr] as Quad::GL_1D evaluates it.
template <class F> double gl10(F&& f, double tl, double tr);

// Basis::product_volume(u, v, du, dv) with 10 GL points.
double product_volume(const Basis1D& U, int ru, const Basis1D& V, int cv, int du, int dv);
// Basis::product_volume_interv(v, w, du, dv, nq) (source/Basis.cpp:83-119):
// integral of u^(du) v^(dv) over the support of w, with nq Gauss-Legendre
// points on each half of the finest of the three supports; 0 unless the three
// supports intersect pairwise.  U, V, W are basis orders of one family.
double product_volume_interv(const Basis1D& B, int ru, int rv, int rw, int du, int dv, int nq);
// Basis::product_edge_dis_v(u, v, sl_u, sl_v, 0, 0, "period").
double product_edge_dis_v_period(const Basis1D& U, int ru, const Basis1D& V, int cv, int slu, int slv);

// Dense level-ordered 1D operator tables, row = "from" (U) basis, column =
// "to" (V) basis, row-major: m[r * cols + c].
struct OperatorTables {
  int rows = 0, cols = 0;
  std::vector<double> u_v, u_vx, ulft_vjp, urgt_vjp, ujp_vjp;
  OperatorTables(const Basis1D& U, const Basis1D& V, bool with_flux = true);
};

// T[alpt_order * n_lagr + lagr_order] = value of Alpert basis at the Lagrange
// point (0 outside the closed support); the transpose that FastLagrIntp builds
// from LagrInterpolation::Lag_pt_Alpt_1D (source/FastMultiplyLU.cpp:1339-1360).
// deriv = 1: the first derivative (LagrInterpolation::eval_Lag_pt_at_Alpt_1D_d1,
// source/Interplation.cpp:104-187), the table of FastLagrIntp::eval_der_up_Lagr.
std::vector<double> alpt_at_lagr_points(const Basis1D& alpt, const Basis1D& lagr, int deriv = 0);
// The same for the Hermite degrees of freedom: value or first derivative of
// the Alpert basis at the Hermite point (HermInterpolation::eval_Her_pt_at_Alpt_1D,
// source/Interplation.cpp:1886-1976).
std::vector<double> alpt_at_herm_points(const Basis1D& alpt, const Basis1D& herm);

// Diagonal of the Lagrange-Lagrange mass matrix (u_v(num,num)), used by the
// interpolation refinement indicator (source/DGAdaptIntp.cpp:236-272).
std::vector<double> lagr_mass_diag(const Basis1D& lagr);

// Hierarchical-surplus data of one 1D Lagrange cell (LagrInterpolation::pwts):
// for cell slot s (level >= 1) the P1 ancestor points (slot, point index) and
// weights wt[p0][ic].  Tables are [slot][ic] and [slot][p0][ic].
struct SurplusTables {
  int p1 = 0, nslots = 0;
  std::vector<int> anc_slot;   // [nslots * p1]
  std::vector<int> anc_point;  // [nslots * p1]
  std::vector<double> wt;      // [nslots * p1 * p1]
  // Lagrange (LagrInterpolation::set_pts_wts_1d_ada_Lag) or Hermite
  // (HermInterpolation::set_pts_wts_1d_ada_Her, source/Interplation.cpp:3166-3316)
  explicit SurplusTables(const Basis1D& intp);
};

// L2 projection of func onto every basis function of B (AllBasis::projection
// with Basis::product_function, include/AllBasis.h:100-108,
// source/Basis.cpp:38-59): 10-point GL over the 16 cells of level 4 for bases
// of level <= 4, over the two halves of the support otherwise.
std::vector<double> project_1d(const Basis1D& B, const std::function<double(double)>& func);

// Entries of the assembled const-coefficient advection operator in one
// dimension: coef*u_vx + coef*(ulft_vjp if coef >= 0 else urgt_vjp), each
// product dropped below 1e-13 as BilinearFormAlpt::assemble_matrix_alpt drops
// triplets (source/BilinearForm.cpp:64-74,691-710).
std::vector<double> linear_operator_table(const OperatorTables& aa, double coef);

// Hash::hash_key arithmetic (int/double mix with truncation kept).
class HashKey {
 public:
  HashKey();
  int key(const int* level, const int* suppt, int dim) const;
  // the level-vector part of the key for dimension dim (1..4): entry
  // sum_t level[t] 16^(dim-1-t); the key is this part plus the support index
  // sum over t of ((suppt[t] - 1) / 2) times the product of 2^level[s], s > t
  const std::vector<int>& level_table(int dim) const;
 private:
  // the reference's arithmetic, term by term
  int key_direct(const int* level, const int* suppt, int dim) const;
  std::vector<std::vector<double>> bicoef_;
  // per dimension: the level-vector part (ind2 + ind3) of the key for every
  // level vector with levels 0..15, computed once with key_direct's arithmetic
  mutable std::vector<int> level_part_[5];
  mutable std::once_flag level_once_[5];
};

// Gauss-Legendre tables of Quad::read_GL_quad_1D (1..10 points).
const std::vector<double>& gl_nodes(int n);
const std::vector<double>& gl_weights(int n);

}  // namespace adg
