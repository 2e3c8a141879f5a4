// PM7 NDDO Hamiltonian for the s/p main-group elements H, C, N, O.
//
// This is the producer of the SCF inner loop's inputs: the one-electron matrix
// H (packed lower triangle), the two-electron integrals W in the order the
// Fock builder consumes them, the core-core repulsion and the isolated-atom
// reference energy.  It runs once per geometry on the host, where the
// reference runs it (hcore before iter), and is not on the per-iteration path.
//
// Methods (published):
//   two-centre integrals  Dewar & Thiel, Theor. Chim. Acta 46, 89 (1977): each
//                         orbital-product charge distribution is a set of
//                         point multipoles (monopole, dipole, quadrupole) made
//                         of point charges, and two distributions interact by
//                         the Klopman-Ohno form q q' / sqrt(r^2 + (rho + rho')^2),
//                         whose additive radii rho reproduce the one-centre
//                         integrals at r = 0.  Evaluated here by one general
//                         point-charge engine for every integral, in a bond
//                         frame whose axis runs from the first atom to the
//                         second.
//   overlaps              Slater-orbital overlaps in prolate spheroidal
//                         coordinates (Mulliken, Rieke, Orloff & Orloff, J.
//                         Chem. Phys. 17, 1248 (1949)): the integrand is
//                         expanded as a polynomial in (xi, eta) and integrated
//                         term by term with the auxiliary integrals A_k, B_k.
//   PM7                   Stewart, J. Mol. Model. 19, 1 (2013): parameters,
//                         core-core function, point-charge switch at 7 A.
//
// Reference behaviour matched (paths relative to the reference root):
//   hcore         src/integrals/hcore.F90:157-288     (H assembly, W order)
//   rotate        src/integrals/rotate.F90:16-124     (per-pair outputs)
//   calpar/inid   src/models/calpar.F90:100-190, src/integrals/mndod.F90:16-61
//   h1elec/diat   src/integrals/h1elec.F90, diat.F90   (resonance block)
//   ccrep         src/integrals/ccrep.F90:16-262      (PM7 core-core)
//   to_point      src/integrals/mndod.F90:3462-3486   (point-charge switch)
//   wstore        src/integrals/mndod.F90:2027-2095   (one-centre block)
//   pdiag         src/moldat.F90:731-790              (diagonal starting density)
//   parameters    src/models/parameters_for_PM7_C.F90, parameters_C.F90:224-231
//   constants     src/conref_C.F90:27-45 (2018 CODATA set, the default)
// The per-pair results equal the reference's routines compiled from its
// sources to rounding (tests/test_pm7_reference.py).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <thread>
#include <atomic>

#include "mopac_hip.h"

namespace {

constexpr double kA0 = 0.529177210903;        // Bohr radius, Angstrom
constexpr double kEV = 27.211386245988;       // Hartree in eV
constexpr double kFPC9 = 23.060547830619029;  // eV -> kcal/mol
constexpr double kSwitchR = 7.0;              // PM7 point-charge switch radius (A)
constexpr double kSwitchK = 0.22;             // and its Gaussian exponent (1/A^2)
constexpr double kPi = 3.14159265358979323846;

// ---- PM7 parameter set (values of the reference's tables, as data) ---------
struct Params {
  int z, ns, np;  // valence s and p electrons of the neutral atom
  double uss, upp, betas, betap, zs, zp;
  double gss, gsp, gpp, gp2, hsp;
  double gk, gl, gm;  // the element's core-core Gaussian: K, L, M
  double dhf_atom;    // experimental heat of formation of the atom (kcal/mol)
};
const Params kPM7[] = {
    {1, 1, 0, -11.070112, 0.0, -8.389745, 0.0, 1.260237, 0.0,
     14.149656, 0.0, 0.0, 0.0, 0.0, 0.177854, 1.428710, 0.991324, 52.102},
    {6, 2, 2, -51.372620, -40.135421, -14.414930, -7.893717, 1.942244, 1.708723,
     12.347323, 11.932801, 10.452226, 9.385498, 0.802634, 0.045888, 5.037055, 1.588715, 170.890},
    {7, 2, 3, -61.623260, -48.949816, -22.110958, -15.464433, 2.354344, 2.028288,
     11.881002, 9.572117, 12.210140, 10.299947, 2.977433, 0.015143, 4.734148, 1.516714, 113.000},
    {8, 2, 4, -96.087736, -71.087738, -67.776339, -20.267981, 5.972309, 2.349017,
     14.955121, 16.088521, 12.403337, 10.499706, 5.028656, -0.016243, 1.871970, 1.844360, 59.559},
};
// Diatomic core-core parameters (alpha_AB in 1/A, x_AB).
struct PairParams {
  int a, b;
  double alpha, x;
};
const PairParams kPM7Pairs[] = {
    {1, 1, 4.051163, 2.845627}, {6, 1, 1.038716, 0.204582}, {6, 6, 2.655746, 0.937818},
    {7, 1, 1.049958, 0.173966}, {7, 6, 2.734636, 0.959163}, {7, 7, 2.786676, 0.863430},
    {8, 1, 1.508030, 0.160602}, {8, 6, 2.850931, 0.826848}, {8, 7, 2.841313, 0.699190},
    {8, 8, 2.457135, 0.356439},
};
// PM7's global core-core terms: a C-C short-range term (scale, exponent) and
// a slow O-H term (scale, exponent).
constexpr double kCCScale = 8.947612, kCCExp = 6.024265;
constexpr double kOHScale = -0.012037, kOHExp = 0.701333;
// The reference writes the r^6 coefficient of its PM6/PM7 pair scale factor
// as a single-precision constant; this is its value as the double it becomes.
const double kR6 = (double)0.0003f;

// ---- the point-charge multipole model -------------------------------------
struct PointCharge {
  double q, x, y, z;  // charge (e) and position (bohr) relative to the atom
};
// One point multipole: up to four charges sharing one additive radius.
struct Multipole {
  PointCharge c[4];
  int n = 0;
  double rho = 0.0;  // additive radius (bohr)
};
// Charge distribution of one product of local orbitals: up to two multipoles
// (the monopole of a p^2 product plus its quadrupole).
struct Distribution {
  Multipole m[2];
  int nm = 0;
  bool monopole = false;  // carries one unit of charge (switches to a point charge)
  bool axial = false;     // the pi x pi' product, fixed by axial symmetry
  int parity = 0;         // bit 0: odd under x -> -x, bit 1: odd under y -> -y
};

// Local orbitals on each atom: 0 = s, 1 = p along the bond axis (sigma, z),
// 2 = p along x (pi), 3 = p along y (pi').  Unordered pairs are indexed
// i(i+1)/2 + j (i >= j), the order MOPAC stores pair blocks in.
inline int pidx(int i, int j) {
  if (i < j) std::swap(i, j);
  return i * (i + 1) / 2 + j;
}

struct Element {
  bool ok = false;
  int z = 0, natorb = 0, nprin = 0;
  double tore = 0.0;
  Params p{};
  double eisol = 0.0;
  double am = 0.0, ad = 0.0, aq = 0.0;  // additive terms as 1/(2 rho) (hartree)
  double d1 = 0.0, d2 = 0.0;            // dipole and quadrupole charge offsets (bohr)
  double rho0 = 0.0, rho1 = 0.0, rho2 = 0.0, rho_core = 0.0;
  Distribution dist[10];
};
Element g_elem[9];
double g_alpb[9][9], g_xfac[9][9];
bool g_init = false;

Multipole monopole(double rho) {
  Multipole m;
  m.c[0] = {1.0, 0, 0, 0};
  m.n = 1;
  m.rho = rho;
  return m;
}
// +1/2 at +d, -1/2 at -d along axis k (0 = x, 1 = y, 2 = z)
Multipole dipole(int k, double d, double rho) {
  Multipole m;
  double a[3] = {0, 0, 0};
  a[k] = d;
  m.c[0] = {0.5, a[0], a[1], a[2]};
  m.c[1] = {-0.5, -a[0], -a[1], -a[2]};
  m.n = 2;
  m.rho = rho;
  return m;
}
// linear quadrupole along axis k: +1/4 at +-2d, -1/2 at the centre
Multipole linear_quadrupole(int k, double d, double rho) {
  Multipole m;
  double a[3] = {0, 0, 0};
  a[k] = 2.0 * d;
  m.c[0] = {0.25, a[0], a[1], a[2]};
  m.c[1] = {0.25, -a[0], -a[1], -a[2]};
  m.c[2] = {-0.5, 0, 0, 0};
  m.n = 3;
  m.rho = rho;
  return m;
}
// square quadrupole in the plane of axes k and l: +1/4 where the two
// coordinates have the same sign, -1/4 where they differ (corners at +-d)
Multipole square_quadrupole(int k, int l, double d, double rho) {
  Multipole m;
  const double s[4][2] = {{1, 1}, {-1, -1}, {1, -1}, {-1, 1}};
  for (int i = 0; i < 4; ++i) {
    double a[3] = {0, 0, 0};
    a[k] = s[i][0] * d;
    a[l] = s[i][1] * d;
    m.c[i] = {0.25 * s[i][0] * s[i][1], a[0], a[1], a[2]};
  }
  m.n = 4;
  m.rho = rho;
  return m;
}

// Interaction energy (hartree) of multipole a at the origin with multipole b
// at (0, 0, r), additive radii summed.
double interact(const Multipole& a, const Multipole& b, double r) {
  const double rho = a.rho + b.rho, rho2 = rho * rho;
  double s = 0.0;
  for (int i = 0; i < a.n; ++i)
    for (int j = 0; j < b.n; ++j) {
      const double dx = a.c[i].x - b.c[j].x, dy = a.c[i].y - b.c[j].y, dz = a.c[i].z - (b.c[j].z + r);
      s += a.c[i].q * b.c[j].q / std::sqrt(dx * dx + dy * dy + dz * dz + rho2);
    }
  return s;
}
// Self-repulsion (hartree) of a set of charges whose pairwise additive term
// is 1/x (that is, each charge's radius is 1/(2x)).
double self_energy(const Multipole& m, double x) {
  const double a2 = 1.0 / (x * x);
  double s = 0.0;
  for (int i = 0; i < m.n; ++i)
    for (int j = 0; j < m.n; ++j) {
      const double dx = m.c[i].x - m.c[j].x, dy = m.c[i].y - m.c[j].y, dz = m.c[i].z - m.c[j].z;
      s += m.c[i].q * m.c[j].q / std::sqrt(dx * dx + dy * dy + dz * dz + a2);
    }
  return s;
}

// Secant iteration x <- x - f(x) (x - x_prev) / (f(x) - f(x_prev)), a fixed
// number of steps.
template <class F>
double secant(F&& f, double x0, double x1, int steps) {
  double f0 = f(x0);
  for (int k = 0; k < steps; ++k) {
    const double f1 = f(x1);
    if (std::fabs(f1 - f0) < 1.0e-25) break;
    const double x2 = x0 + (x1 - x0) * (0.0 - f0) / (f1 - f0);
    x0 = x1;
    f0 = f1;
    x1 = x2;
  }
  return x1;
}

// Multipole geometry and additive radii of one element (Dewar & Thiel):
//  d1   the dipole length of the s.p_sigma distribution, <ns|z|np_z>, from
//       the Slater radial integral N_s N_p (2n+1)! / (zs+zp)^(2n+2) / sqrt(3);
//  d2   sqrt(<np|r^2|np> / 5), <r^2> = (2n+1)(2n+2) / (2 zp)^2, the
//       quadrupole offset (linear quadrupoles put +1/4 at +-2 d2);
//  rho  chosen so each multipole's self-repulsion equals the one-centre
//       integral it stands for: Gss (monopole), Hsp (dipole), and
//       Hpp = (Gpp - Gp2)/2 (the square quadrupole of p_x p_y).
// The dipole and quadrupole radii are solved by five secant steps from the
// small-x limits of the self-energies (D^2 x^3 and 3 D^4 x^5) and that start
// + 0.04, the steps MOPAC's calpar takes: the quadrupole term is then ~1e-9
// short of the exact root, and parity with the reference's integrals needs
// the same five steps.
void derive(Element& e) {
  const Params& p = e.p;
  e.z = p.z;
  e.tore = p.ns + p.np;
  e.natorb = (p.zp > 0.0) ? 4 : 1;
  e.nprin = (p.z <= 2) ? 1 : 2;
  const double ev = kEV;
  if (e.natorb == 1) {
    e.eisol = p.uss;
    e.am = p.gss / ev;
    e.ad = e.am;
    e.aq = e.am;
  } else {
    const double zp = std::max(0.3, p.zp), zs = p.zs;
    const double hsp = std::max(1.0e-7, p.hsp);
    const double hpp = std::max(0.1, 0.5 * (p.gpp - p.gp2));
    // Isolated-atom electronic energy of the valence configuration s^ns p^np
    // (Hund's rule): one-electron terms, then the Coulomb and exchange pairs.
    const int ns = p.ns, np = p.np;
    const int unpaired_pairs = std::min(np, 6 - np);
    const double n_ss = std::max(ns - 1, 0);
    const double n_sp = ns * np;
    const double n_sp_exch = -0.5 * ns * np;
    const double n_p2 = (np * (np - 1)) / 2 + 0.5 * (unpaired_pairs * (unpaired_pairs - 1)) / 2;
    const double n_pp = -0.5 * (unpaired_pairs * (unpaired_pairs - 1)) / 2;
    e.eisol = p.uss * ns + p.upp * np + p.gss * n_ss + p.gpp * n_pp + p.gsp * n_sp + p.gp2 * n_p2 +
              hsp * n_sp_exch;
    const int n = e.nprin;
    const double fact2n = (n == 1) ? 2.0 : 24.0, fact2n1 = (n == 1) ? 6.0 : 120.0;
    const double ns_norm = std::pow(2.0 * zs, n + 0.5) / std::sqrt(fact2n);
    const double np_norm = std::pow(2.0 * zp, n + 0.5) / std::sqrt(fact2n);
    e.d1 = ns_norm * np_norm * fact2n1 / std::pow(zs + zp, 2 * n + 2) / std::sqrt(3.0);
    e.d2 = std::sqrt((2.0 * n + 1.0) * (2.0 * n + 2.0) / (4.0 * zp * zp) / 5.0);
    const Multipole dip = dipole(2, e.d1, 0.0);
    const Multipole quad = square_quadrupole(0, 1, e.d2, 0.0);
    const double x_dip = std::cbrt(hsp / ev / (e.d1 * e.d1));
    e.ad = secant([&](double x) { return self_energy(dip, x) - hsp / ev; }, x_dip, x_dip + 0.04, 5);
    const double x_quad = std::pow(hpp / ev / (3.0 * std::pow(e.d2, 4)), 0.2);
    e.aq = secant([&](double x) { return self_energy(quad, x) - hpp / ev; }, x_quad, x_quad + 0.04, 5);
    e.am = p.gss / ev;
  }
  e.rho0 = 0.5 / e.am;
  e.rho1 = (e.natorb > 1) ? 0.5 / e.ad : 0.0;
  e.rho2 = (e.natorb > 1) ? 0.5 / e.aq : 0.0;
  e.rho_core = e.rho0;  // PM7 gives H, C, N, O no separate core radius
  // the ten orbital-product distributions
  for (auto& d : e.dist) d = Distribution{};
  auto set = [&](int a, int b, std::initializer_list<Multipole> ms, bool mono, int parity) {
    Distribution& d = e.dist[pidx(a, b)];
    for (const Multipole& m : ms) d.m[d.nm++] = m;
    d.monopole = mono;
    d.parity = parity;
  };
  set(0, 0, {monopole(e.rho0)}, true, 0);
  if (e.natorb > 1) {
    set(1, 0, {dipole(2, e.d1, e.rho1)}, false, 0);
    set(2, 0, {dipole(0, e.d1, e.rho1)}, false, 1);
    set(3, 0, {dipole(1, e.d1, e.rho1)}, false, 2);
    set(1, 1, {monopole(e.rho0), linear_quadrupole(2, e.d2, e.rho2)}, true, 0);
    set(2, 2, {monopole(e.rho0), linear_quadrupole(0, e.d2, e.rho2)}, true, 0);
    set(3, 3, {monopole(e.rho0), linear_quadrupole(1, e.d2, e.rho2)}, true, 0);
    set(2, 1, {square_quadrupole(0, 2, e.d2, e.rho2)}, false, 1);
    set(3, 1, {square_quadrupole(1, 2, e.d2, e.rho2)}, false, 2);
    e.dist[pidx(3, 2)].axial = true;
    e.dist[pidx(3, 2)].parity = 3;
  }
  e.ok = true;
}

void init_params() {
  if (g_init) return;
  for (const Params& p : kPM7) {
    g_elem[p.z].p = p;
    derive(g_elem[p.z]);
  }
  std::memset(g_alpb, 0, sizeof(g_alpb));
  std::memset(g_xfac, 0, sizeof(g_xfac));
  for (const PairParams& q : kPM7Pairs) {
    g_alpb[q.a][q.b] = g_alpb[q.b][q.a] = q.alpha;
    g_xfac[q.a][q.b] = g_xfac[q.b][q.a] = q.x;
  }
  g_init = true;
}

inline bool supported(int z) { return z == 1 || z == 6 || z == 7 || z == 8; }

// PM7's switch from NDDO to point-charge electrostatics (to_point): the
// weight c of the NDDO value at r (Angstrom) and the point-charge e-e value
// 1/r (eV).
inline double switch_weight(double r_ang, double& point_ev) {
  point_ev = kEV * kA0 / r_ang;
  return (r_ang < kSwitchR) ? 1.0 - std::exp(-(r_ang - kSwitchR) * (r_ang - kSwitchR) * kSwitchK) : 0.0;
}

// Two-centre repulsion (eV) between distribution a of atom A (origin) and
// distribution b of atom B at (0, 0, r bohr), before the point-charge switch.
double repulsion(const Distribution& a, const Distribution& b, double r) {
  double s = 0.0;
  for (int i = 0; i < a.nm; ++i)
    for (int j = 0; j < b.nm; ++j) s += interact(a.m[i], b.m[j], r);
  return kEV * s;
}

// Bond-frame tensors of one atom pair: local two-electron integrals L over
// unordered local pairs (A's pair down, B's across), electron-core
// attractions of each atom's distributions in the other's core field, and
// the core-core Coulomb function gab, all in eV with the point-charge switch
// applied.  r_ang: distance in Angstrom.
struct LocalPair {
  double L[10][10];
  double core_a[10], core_b[10];
  double gab;
};
void local_pair(const Element& A, const Element& B, double r_ang, LocalPair& lp) {
  std::memset(&lp, 0, sizeof(lp));
  const double r = r_ang / kA0;
  double point;
  const double c = switch_weight(r_ang, point);
  const int la = A.natorb * (A.natorb + 1) / 2, lb = B.natorb * (B.natorb + 1) / 2;
  double xx_xx = 0.0, xx_yy = 0.0;
  for (int ab = 0; ab < la; ++ab)
    for (int cd = 0; cd < lb; ++cd) {
      const Distribution& da = A.dist[ab];
      const Distribution& db = B.dist[cd];
      if (da.parity != db.parity || da.axial != db.axial) continue;
      if (da.axial) continue;  // filled below
      const double v = repulsion(da, db, r);
      lp.L[ab][cd] = da.monopole && db.monopole ? v * c + (1.0 - c) * point : v * c;
    }
  if (A.natorb > 1 && B.natorb > 1) {
    // (pi pi'|pi pi') = [(pi pi|pi pi) - (pi pi|pi' pi')] / 2 about the bond axis
    xx_xx = lp.L[pidx(2, 2)][pidx(2, 2)];
    xx_yy = lp.L[pidx(2, 2)][pidx(3, 3)];
    lp.L[pidx(3, 2)][pidx(3, 2)] = 0.5 * (xx_xx - xx_yy);
  }
  // electron-core attraction: each core is a unit monopole of radius rho_core
  const Multipole core_of_a = monopole(A.rho_core), core_of_b = monopole(B.rho_core);
  const double point_b = -(kEV / r) * B.tore, point_a = -(kEV / r) * A.tore;
  for (int ab = 0; ab < la; ++ab) {
    const Distribution& d = A.dist[ab];
    if (d.parity != 0 || d.axial || d.nm == 0) continue;
    double s = 0.0;
    for (int i = 0; i < d.nm; ++i) s += interact(d.m[i], core_of_b, r);
    const double v = -B.tore * kEV * s;
    lp.core_a[ab] = d.monopole ? v * c + (1.0 - c) * point_b : v * c;
  }
  for (int cd = 0; cd < lb; ++cd) {
    const Distribution& d = B.dist[cd];
    if (d.parity != 0 || d.axial || d.nm == 0) continue;
    double s = 0.0;
    for (int i = 0; i < d.nm; ++i) s += interact(core_of_a, d.m[i], r);
    const double v = -A.tore * kEV * s;
    lp.core_b[cd] = d.monopole ? v * c + (1.0 - c) * point_a : v * c;
  }
  const double rc = A.rho_core + B.rho_core;
  const double gab = kEV / std::sqrt(r * r + rc * rc);
  lp.gab = gab * c + (1.0 - c) * point;
}

// PM7 core-core repulsion (eV) at r bohr, given the switched gab: the
// screened Coulomb term times the diatomic scale factor, the elements'
// Gaussians, and the short-range r^-12 wall.
double core_core(const Element& A, const Element& B, double r_bohr, double gab) {
  const double r = r_bohr * kA0;
  const int za = A.z, zb = B.z, hi = std::max(za, zb), lo = std::min(za, zb);
  const double zz = A.tore * B.tore;
  const double x = g_xfac[za][zb];
  const double alpha = g_alpb[za][zb] < 1e-6 ? 1.2 : g_alpb[za][zb];
  double scale;
  if (lo == 1 && (hi == 6 || hi == 7)) {
    scale = 1.0 + 2.0 * x * std::exp(-alpha * r * r);
  } else if (lo == 1 && hi == 8) {
    scale = 1.0 + 2.0 * x * std::exp(-alpha * r * r) - kOHScale * std::exp(-kOHExp * r * 2);
  } else {
    scale = 1.0 + 2.0 * x * std::exp(-alpha * (r + kR6 * std::pow(r, 6)));
    if (lo == 6 && hi == 6) scale += kCCScale * std::exp(-kCCExp * r);
  }
  double e = zz * gab * scale;
  double gauss = 0.0;
  for (const Element* E : {&A, &B}) {
    const double ax = E->p.gl * (r - E->p.gm) * (r - E->p.gm);
    if (ax < 25.0) gauss += zz / r * E->p.gk * std::exp(-ax);
  }
  e += gauss;
  const double rr = r / (std::pow((double)za, 0.3333) + std::pow((double)zb, 0.3333));
  if (rr < 3.0) e += std::min(1.0e-8 / std::pow(rr, 12), 1.0e5);
  return e;
}

// ---- Slater overlaps in prolate spheroidal coordinates ----------------------
// r_A = (R/2)(xi + eta), r_B = (R/2)(xi - eta), z_A = (R/2)(1 + xi eta),
// z_B = (R/2)(xi eta - 1) (z along A -> B), rho^2 = (R/2)^2 (xi^2-1)(1-eta^2),
// dV = (R/2)^3 (xi^2 - eta^2) dxi deta dphi.  Polynomials in (xi, eta) carry
// the (R/2) powers separately.
struct Poly {
  static constexpr int N = 9;
  double c[N][N];  // c[i][j] xi^i eta^j
  Poly() { std::memset(c, 0, sizeof(c)); }
};
Poly poly_mul(const Poly& a, const Poly& b) {
  Poly r;
  for (int i = 0; i < Poly::N; ++i)
    for (int j = 0; j < Poly::N; ++j) {
      if (a.c[i][j] == 0.0) continue;
      for (int k = 0; k + i < Poly::N; ++k)
        for (int l = 0; l + j < Poly::N; ++l) r.c[i + k][j + l] += a.c[i][j] * b.c[k][l];
    }
  return r;
}
Poly poly_of(std::initializer_list<std::initializer_list<double>> terms) {
  // each term {coefficient, xi power, eta power}
  Poly p;
  for (auto t : terms) {
    auto it = t.begin();
    const double c = *it++;
    const int i = (int)*it++, j = (int)*it;
    p.c[i][j] += c;
  }
  return p;
}

// A_k(p) = int_1^inf xi^k e^(-p xi) dxi, k = 0..kmax (p > 0): the closed
// recurrence A_k = (k A_{k-1} + e^-p) / p, stable upward.
void aux_a(double p, int kmax, double* a) {
  const double e = std::exp(-p);
  a[0] = e / p;
  for (int k = 1; k <= kmax; ++k) a[k] = (k * a[k - 1] + e) / p;
}
// B_k(t) = int_-1^1 eta^k e^(-t eta) deta, k = 0..kmax.  Upward recurrence
// B_k = (k B_{k-1} + (-1)^k e^t - e^-t) / t where it is stable, otherwise
// the power series in t, whose length for each |t| band is the one the
// reference's overlap code uses (src/integrals/set.F90, bintgs): its overlaps
// carry that series' truncation (up to ~1e-7 relative for |t| near 0.5), and
// parity needs the same sums.
void aux_b(double t, int kmax, double* b) {
  const double at = std::fabs(t);
  // (|t| upper bound, largest k served by the recurrence, series terms beyond it)
  struct Band {
    double tmax;
    int kmax_rec, last;
  };
  static const Band bands[] = {{0.5, -1, 6}, {1.0, 5, 7}, {2.0, 7, 12}, {3.0, 10, 15}};
  int kmax_rec = 1 << 20, last = 0;
  if (at <= 1.0e-6) {
    for (int k = 0; k <= kmax; ++k) b[k] = (k % 2 == 0) ? 2.0 / (k + 1.0) : 0.0;
    return;
  }
  for (const Band& band : bands)
    if (at <= band.tmax) {
      kmax_rec = band.kmax_rec;
      last = band.last;
      break;
    }
  if (kmax <= kmax_rec) {
    const double ep = std::exp(t), em = 1.0 / ep;
    b[0] = (ep - em) / t;
    for (int k = 1; k <= kmax; ++k) b[k] = (k * b[k - 1] + ((k % 2) ? -ep : ep) - em) / t;
    return;
  }
  for (int k = 0; k <= kmax; ++k) {
    // sum_{m=0..last} (-t)^m / m! * int eta^(k+m) (zero for odd k + m)
    double s = 0.0, term = 1.0;
    for (int m = 0; m <= last; ++m) {
      if (m > 0) term *= -t / m;
      if ((k + m) % 2 == 0) s += term * 2.0 / (k + m + 1.0);
    }
    b[k] = s;
  }
}

// Overlap of two Slater orbitals in the bond frame: orbital a on A (principal
// number na, l = la, zeta za), orbital b on B; m = 0 (sigma) or 1 (pi, both
// along the same perpendicular axis).  R in bohr.
double sto_overlap(int na, int la, double za, int nb, int lb, double zb, int m, double R) {
  const Poly rA = poly_of({{1, 1, 0}, {1, 0, 1}});   // xi + eta
  const Poly rB = poly_of({{1, 1, 0}, {-1, 0, 1}});  // xi - eta
  const Poly zA = poly_of({{1, 0, 0}, {1, 1, 1}});   // 1 + xi eta
  const Poly zB = poly_of({{-1, 0, 0}, {1, 1, 1}});  // xi eta - 1
  const Poly rho2 = poly_of({{-1, 0, 0}, {1, 2, 0}, {1, 0, 2}, {-1, 2, 2}});  // (xi^2-1)(1-eta^2)
  Poly f = poly_of({{1, 2, 0}, {-1, 0, 2}});  // xi^2 - eta^2
  int rpow = 3;
  for (int k = 0; k < na - 1 - la; ++k) f = poly_mul(f, rA), ++rpow;
  for (int k = 0; k < nb - 1 - lb; ++k) f = poly_mul(f, rB), ++rpow;
  double ang;  // angular normalization times the phi integral
  if (m == 0) {
    if (la == 1) f = poly_mul(f, zA), ++rpow;
    if (lb == 1) f = poly_mul(f, zB), ++rpow;
    ang = 2.0 * kPi * (la == 1 ? std::sqrt(3.0 / (4.0 * kPi)) : 1.0 / std::sqrt(4.0 * kPi)) *
          (lb == 1 ? std::sqrt(3.0 / (4.0 * kPi)) : 1.0 / std::sqrt(4.0 * kPi));
  } else {
    f = poly_mul(f, rho2);
    rpow += 2;
    ang = kPi * 3.0 / (4.0 * kPi);
  }
  auto fact = [](int n) { double r = 1.0; for (int i = 2; i <= n; ++i) r *= i; return r; };
  const double norm = std::pow(2.0 * za, na + 0.5) / std::sqrt(fact(2 * na)) *
                      std::pow(2.0 * zb, nb + 0.5) / std::sqrt(fact(2 * nb));
  // the auxiliary integrals up to the polynomial's own degrees (the band
  // choice in aux_b depends on the highest order asked for)
  int imax = 0, jmax = 0;
  for (int i = 0; i < Poly::N; ++i)
    for (int j = 0; j < Poly::N; ++j)
      if (f.c[i][j] != 0.0) imax = std::max(imax, i), jmax = std::max(jmax, j);
  const double p = 0.5 * R * (za + zb), t = 0.5 * R * (za - zb);
  double A[Poly::N], B[Poly::N];
  aux_a(p, imax, A);
  aux_b(t, jmax, B);
  double s = 0.0;
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j)
      if (f.c[i][j] != 0.0) s += f.c[i][j] * A[i] * B[j];
  return norm * ang * std::pow(0.5 * R, rpow) * s;
}

// Bond frame of the pair A -> B: rows e[0] (sigma, toward B), e[1] (pi),
// e[2] (pi'), a right-handed orthonormal set.  d = xB - xA.
void bond_frame(const double d[3], double r, double e[3][3]) {
  for (int k = 0; k < 3; ++k) e[0][k] = d[k] / r;
  // pi: the coordinate axis least aligned with the bond, made orthogonal to it
  int ax = 0;
  for (int k = 1; k < 3; ++k)
    if (std::fabs(e[0][k]) < std::fabs(e[0][ax])) ax = k;
  double v[3] = {0, 0, 0};
  v[ax] = 1.0;
  const double dot = e[0][ax];
  for (int k = 0; k < 3; ++k) v[k] -= dot * e[0][k];
  const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) e[1][k] = v[k] / nv;
  e[2][0] = e[0][1] * e[1][2] - e[0][2] * e[1][1];
  e[2][1] = e[0][2] * e[1][0] - e[0][0] * e[1][2];
  e[2][2] = e[0][0] * e[1][1] - e[0][1] * e[1][0];
}

struct PairOut {
  double w[100];
  double e1b[10], e2a[10];
  double di[4][4];
  double enuc;
};

// Everything one atom pair contributes: W block, both core-attraction blocks,
// the resonance block and the core-core term (rotate + h1elec).  Atom A has
// the larger index (hcore's i).  The resonance block is zero when r^2 >
// res_cut_sq (h1elec's CUTOFS test) or r^2 >= ovl_cut_sq (diat's overlap
// cutoff); hcore uses 1e4 and 225 (15 A), hcore_for_MOZYME the 7 A CUTOFS and
// fillij's CUTOF1.
void pair_integrals(const Element& A, const Element& B, const double* xa, const double* xb, PairOut& out,
                    double res_cut_sq = 1.0e4, double ovl_cut_sq = 225.0) {
  const int li = A.natorb, lj = B.natorb;
  const int limij = li * (li + 1) / 2, limkl = lj * (lj + 1) / 2;
  std::memset(&out, 0, sizeof(out));
  const double d[3] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]};
  const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (rsq < 0.00002) return;  // coincident atoms: rotate returns zeros
  const double r_ang = std::sqrt(rsq), r = r_ang / kA0;
  double e[3][3];
  bond_frame(d, r_ang, e);
  // U[mu][a]: molecular orbital mu (s, x, y, z) in the local orbitals a
  // (s, sigma, pi, pi') of the same atom.
  double U[4][4] = {{1, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) U[1 + k][1 + a] = e[a][k];
  // Pair transform T[molecular pair][local pair] (an off-diagonal local pair
  // collects both orders).
  auto build_T = [&](int norb, double T[10][10]) {
    std::memset(T, 0, sizeof(double) * 100);
    for (int mu = 0; mu < norb; ++mu)
      for (int nu = 0; nu <= mu; ++nu)
        for (int a = 0; a < norb; ++a)
          for (int b = 0; b < norb; ++b) T[pidx(mu, nu)][pidx(a, b)] += U[mu][a] * U[nu][b];
  };
  double TA[10][10], TB[10][10];
  build_T(li, TA);
  build_T(lj, TB);
  // Resonance block: overlap in the molecular frame times (beta_a + beta_b)/2.
  if (rsq <= res_cut_sq && rsq < ovl_cut_sq) {
    const int nA = A.nprin, nB = B.nprin;
    double S[4][4] = {};  // local: [a on A][b on B]
    S[0][0] = sto_overlap(nA, 0, A.p.zs, nB, 0, B.p.zs, 0, r);
    if (lj > 1) S[0][1] = sto_overlap(nA, 0, A.p.zs, nB, 1, B.p.zp, 0, r);
    if (li > 1) S[1][0] = sto_overlap(nA, 1, A.p.zp, nB, 0, B.p.zs, 0, r);
    if (li > 1 && lj > 1) {
      S[1][1] = sto_overlap(nA, 1, A.p.zp, nB, 1, B.p.zp, 0, r);
      S[2][2] = S[3][3] = sto_overlap(nA, 1, A.p.zp, nB, 1, B.p.zp, 1, r);
    }
    const double ba[4] = {0.5 * A.p.betas, 0.5 * A.p.betap, 0.5 * A.p.betap, 0.5 * A.p.betap};
    const double bb[4] = {0.5 * B.p.betas, 0.5 * B.p.betap, 0.5 * B.p.betap, 0.5 * B.p.betap};
    for (int mu = 0; mu < li; ++mu)
      for (int nu = 0; nu < lj; ++nu) {
        double s = 0.0;
        for (int a = 0; a < li; ++a)
          for (int b = 0; b < lj; ++b) s += U[mu][a] * U[nu][b] * S[a][b];
        out.di[mu][nu] = s * (ba[mu] + bb[nu]);
      }
  }
  LocalPair lp;
  local_pair(A, B, r_ang, lp);
  // (mu nu | lam sig) = sum_ab,cd TA[mn][ab] L[ab][cd] TB[ls][cd]
  double M[10][10];
  for (int mn = 0; mn < limij; ++mn)
    for (int cd = 0; cd < limkl; ++cd) {
      double acc = 0.0;
      for (int ab = 0; ab < limij; ++ab) acc += TA[mn][ab] * lp.L[ab][cd];
      M[mn][cd] = acc;
    }
  for (int mn = 0; mn < limij; ++mn)
    for (int ls = 0; ls < limkl; ++ls) {
      double acc = 0.0;
      for (int cd = 0; cd < limkl; ++cd) acc += M[mn][cd] * TB[ls][cd];
      out.w[mn * limkl + ls] = acc;
    }
  for (int mn = 0; mn < limij; ++mn) {
    double acc = 0.0;
    for (int ab = 0; ab < limij; ++ab) acc += TA[mn][ab] * lp.core_a[ab];
    out.e1b[mn] = acc;
  }
  for (int ls = 0; ls < limkl; ++ls) {
    double acc = 0.0;
    for (int cd = 0; cd < limkl; ++cd) acc += TB[ls][cd] * lp.core_b[cd];
    out.e2a[ls] = acc;
  }
  out.enuc = core_core(A, B, r, lp.gab);
}

// Value of the (ss|ss), (s sigma|ss) and (ss|s sigma) repulsions of a pair
// (eV, switched), sigma pointing from A toward B; for MOZYME's outer2 shell.
void pair_dipole_terms(const Element& A, const Element& B, double r_ang, double& ss, double& sp_a,
                       double& sp_b) {
  double point;
  const double c = switch_weight(r_ang, point);
  const double r = r_ang / kA0;
  ss = repulsion(A.dist[0], B.dist[0], r) * c + (1.0 - c) * point;
  sp_a = (A.natorb > 1) ? repulsion(A.dist[pidx(1, 0)], B.dist[0], r) * c : 0.0;
  sp_b = (B.natorb > 1) ? repulsion(A.dist[0], B.dist[pidx(1, 0)], r) * c : 0.0;
}

void one_center_block(const Element& E, double* w) {
  const int ilim = E.natorb * (E.natorb + 1) / 2;
  for (int i = 0; i < ilim * ilim; ++i) w[i] = 0.0;
  auto W = [&](int i, int j) -> double& { return w[(j - 1) * ilim + (i - 1)]; };
  const Params& p = E.p;
  W(1, 1) = p.gss;
  if (E.natorb > 2) {
    const int ipx = 3, ipy = 6, ipz = 10;
    W(ipx, 1) = W(ipy, 1) = W(ipz, 1) = p.gsp;
    W(1, ipx) = W(1, ipy) = W(1, ipz) = p.gsp;
    W(ipx, ipx) = W(ipy, ipy) = W(ipz, ipz) = p.gpp;
    W(ipy, ipx) = W(ipz, ipx) = W(ipz, ipy) = p.gp2;
    W(ipx, ipy) = W(ipx, ipz) = W(ipy, ipz) = p.gp2;
    W(2, 2) = W(4, 4) = W(7, 7) = std::max(1.0e-7, p.hsp);
    W(5, 5) = W(8, 8) = W(9, 9) = 0.5 * (p.gpp - p.gp2);
  }
}

}  // namespace

extern "C" {

int mhip_pm7_sizes(int numat, const int* nat, int* norbs, long long* n2elec) {
  init_params();
  long long n = 0, nw = 0;
  std::vector<int> no(numat);
  for (int i = 0; i < numat; ++i) {
    if (nat[i] < 0 || nat[i] > 8 || !supported(nat[i])) return -1;
    no[i] = g_elem[nat[i]].natorb;
    n += no[i];
  }
  for (int i = 0; i < numat; ++i) {
    const int li = no[i] * (no[i] + 1) / 2;
    for (int j = 0; j < i; ++j) nw += (long long)li * (no[j] * (no[j] + 1) / 2);
    nw += (long long)li * li;
  }
  *norbs = (int)n;
  *n2elec = nw;
  return 0;
}

int mhip_pm7_hamiltonian(int numat, const int* nat, const double* coord, int charge,
                         int* nfirst, int* nlast, double* h, double* w, double* enuclr,
                         double* atheat, double* pdiag) {
  init_params();
  int norbs = 0;
  long long nw = 0;
  if (mhip_pm7_sizes(numat, nat, &norbs, &nw) != 0) return -1;
  // orbital ranges (1-based, as the reference stores them)
  int k = 1;
  for (int i = 0; i < numat; ++i) {
    nfirst[i] = k;
    k += g_elem[nat[i]].natorb;
    nlast[i] = k - 1;
  }
  // W offsets in hcore order: for atom i, pairs (i, j<i) then the one-centre block
  std::vector<long long> off(numat + 1);
  long long kr = 0;
  for (int i = 0; i < numat; ++i) {
    off[i] = kr;
    const int li = g_elem[nat[i]].natorb * (g_elem[nat[i]].natorb + 1) / 2;
    for (int j = 0; j < i; ++j) {
      const int lj = g_elem[nat[j]].natorb * (g_elem[nat[j]].natorb + 1) / 2;
      kr += (long long)li * lj;
    }
    kr += (long long)li * li;
  }
  off[numat] = kr;
  const long long mpack = (long long)norbs * (norbs + 1) / 2;
  for (long long q = 0; q < mpack; ++q) h[q] = 0.0;
  // diagonal one-centre terms
  for (int i = 0; i < numat; ++i) {
    const Element& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a) {
      const long long o = nfirst[i] - 1 + a;
      h[o * (o + 1) / 2 + o] = (a == 0) ? E.p.uss : E.p.upp;
    }
  }
  // Per-atom accumulation of core-attraction blocks, done per row i in pair
  // order j = 0..i-1 (same summation order as the reference); the e2a blocks
  // land on atom j and are reduced afterwards in that same (i, j) order.
  std::vector<double> e2a_acc((size_t)numat * 10, 0.0);
  std::vector<double> e1b_acc((size_t)numat * 10, 0.0);
  std::vector<double> enuc_row(numat, 0.0);
  // The j-side accumulation is order-sensitive; store per-pair e2a for a
  // deterministic serial reduction.
  std::vector<double> e2a_pair;
  const long long npairs = (long long)numat * (numat - 1) / 2;
  e2a_pair.assign((size_t)npairs * 10, 0.0);
  // Rows are independent (disjoint W ranges, disjoint H off-diagonal blocks);
  // hand them out dynamically to a few host threads.
  std::atomic<int> next_row{0};
  auto worker = [&]() {
  for (int i = next_row++; i < numat; i = next_row++) {
    const Element& A = g_elem[nat[i]];
    const int li = A.natorb, limij = li * (li + 1) / 2;
    long long kk = off[i];
    double rowsum = 0.0;
    for (int j = 0; j < i; ++j) {
      const Element& B = g_elem[nat[j]];
      const int lj = B.natorb, limkl = lj * (lj + 1) / 2;
      PairOut po;
      pair_integrals(A, B, coord + 3 * i, coord + 3 * j, po);
      for (int q = 0; q < limij * limkl; ++q) w[kk + q] = po.w[q];
      kk += (long long)limij * limkl;
      rowsum += po.enuc;
      // resonance block: rows on atom i, columns on atom j (j precedes i)
      for (int a = 0; a < li; ++a) {
        const long long row = nfirst[i] - 1 + a;
        for (int b = 0; b < lj; ++b) {
          const long long col = nfirst[j] - 1 + b;
          h[row * (row + 1) / 2 + col] += po.di[a][b];
        }
      }
      for (int q = 0; q < limij; ++q) e1b_acc[(size_t)i * 10 + q] += po.e1b[q];
      const long long pid = (long long)i * (i - 1) / 2 + j;
      for (int q = 0; q < limkl; ++q) e2a_pair[(size_t)pid * 10 + q] = po.e2a[q];
    }
    one_center_block(A, w + kk);
    enuc_row[i] = rowsum;
  }
  };
  unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (numat < 64) nthr = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nthr; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  // Core-attraction blocks onto the diagonal atom blocks, in hcore order.
  double en = 0.0;
  for (int i = 0; i < numat; ++i) {
    en += enuc_row[i];
    const int li = g_elem[nat[i]].natorb;
    for (int a = 0; a < li; ++a)
      for (int b = 0; b <= a; ++b) {
        const long long row = nfirst[i] - 1 + a, col = nfirst[i] - 1 + b;
        h[row * (row + 1) / 2 + col] += e1b_acc[(size_t)i * 10 + pidx(a, b)];
      }
  }
  for (int i = 0; i < numat; ++i)
    for (int j = 0; j < i; ++j) {
      const long long pid = (long long)i * (i - 1) / 2 + j;
      const int lj = g_elem[nat[j]].natorb;
      for (int a = 0; a < lj; ++a)
        for (int b = 0; b <= a; ++b) {
          const long long row = nfirst[j] - 1 + a, col = nfirst[j] - 1 + b;
          h[row * (row + 1) / 2 + col] += e2a_pair[(size_t)pid * 10 + pidx(a, b)];
        }
    }
  *enuclr = en;
  // Isolated-atom reference: sum of atomic heats minus electronic energies.
  double eheat = 0.0, eisol = 0.0;
  for (int i = 0; i < numat; ++i) {
    eheat += g_elem[nat[i]].p.dhf_atom;
    eisol += g_elem[nat[i]].eisol;
  }
  *atheat = eheat - eisol * kFPC9;
  // Diagonal starting density
  const double yy = (double)charge / (norbs + 1.0e-10);
  for (int i = 0; i < numat; ++i) {
    const Element& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a)
      pdiag[nfirst[i] - 1 + a] = (E.natorb == 1) ? E.tore - yy : E.tore * 0.25 - yy;
  }
  return 0;
}

double mhip_ev_to_kcal(void) { return kFPC9; }

// The one-electron matrix on the MOZYME sparse storage and the core-core
// energy, as hcore_for_MOZYME forms them for a discrete system
// (src/MOZYME/hcore_for_MOZYME.F90:100-260): stored pairs (nijbo >= 0) take
// the NDDO terms (electron-core attraction, core-core, and the resonance
// block up to overlap_cutoff = CUTOFS, fillij.F90:120-126, and below
// dipole_cutoff = CUTOF1, diat's overlap limit under MOZYME);
// pairs of the dipole shell (-2) outer2's point charge plus dipole
// (src/MOZYME/outer2.F90:16-150, from the NDDO (ss|ss), (s sigma|ss) and
// (ss|s sigma) integrals); the others outer1's point charge
// (src/MOZYME/outer1.F90:16-72) with PM7's switch to the bare Coulomb.
int mhip_pm7_hamiltonian_mozyme(int numat, const int* nat, const double* coord, const int* nijbo,
                                long long mpack, double overlap_cutoff, double dipole_cutoff, double* h,
                                double* enuclr) {
  init_params();
  if (!(overlap_cutoff > 0.0) || !(dipole_cutoff > 0.0)) return -1;
  const double res_sq = overlap_cutoff * overlap_cutoff, ovl_sq = dipole_cutoff * dipole_cutoff;
  int norbs = 0;
  long long nw = 0;
  if (mhip_pm7_sizes(numat, nat, &norbs, &nw) != 0) return -1;
  auto nij = [&](int a, int b) { return nijbo[a + (size_t)b * numat]; };
  for (int i = 0; i < numat; ++i) {
    const int no = g_elem[nat[i]].natorb;
    if (nij(i, i) < 0 || nij(i, i) + no * (no + 1) / 2 > mpack) return -1;
    for (int j = 0; j < i; ++j)
      if (nij(i, j) >= 0 && nij(i, j) + no * g_elem[nat[j]].natorb > mpack) return -1;
  }
  for (long long q = 0; q < mpack; ++q) h[q] = 0.0;
  for (int i = 0; i < numat; ++i) {
    const Element& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a) h[nij(i, i) + a * (a + 1) / 2 + a] = (a == 0) ? E.p.uss : E.p.upp;
  }
  const long long npairs = (long long)numat * (numat - 1) / 2;
  std::vector<double> e1b_acc((size_t)numat * 10, 0.0), enuc_row(numat, 0.0);
  std::vector<double> e2a_pair((size_t)npairs * 10, 0.0);
  std::atomic<int> next_row{0};
  auto worker = [&]() {
    for (int i = next_row++; i < numat; i = next_row++) {
      const Element& A = g_elem[nat[i]];
      const int li = A.natorb;
      double rowsum = 0.0;
      for (int j = 0; j < i; ++j) {
        const Element& B = g_elem[nat[j]];
        const int lj = B.natorb;
        double e1b[10] = {0}, e2a[10] = {0}, enuc = 0.0;
        const int o = nij(i, j);
        if (o >= 0) {
          PairOut po;
          pair_integrals(A, B, coord + 3 * i, coord + 3 * j, po, res_sq, ovl_sq);
          for (int a = 0; a < li; ++a)
            for (int b = 0; b < lj; ++b) h[o + a * lj + b] += po.di[a][b];
          std::memcpy(e1b, po.e1b, sizeof(e1b));
          std::memcpy(e2a, po.e2a, sizeof(e2a));
          enuc = po.enuc;
        } else {
          const double* xi = coord + 3 * i;
          const double* xj = coord + 3 * j;
          double u[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};  // i -> j
          const double rij = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
          double w1, sp_i, sp_j;
          pair_dipole_terms(A, B, rij, w1, sp_i, sp_j);
          if (o == -2) {
            // outer2: the (p, s) blocks of each atom in the other's charge
            // take the (s p_sigma|ss) repulsions, rotated to the molecular
            // frame by the bond's direction cosines (the z cosine held at +-1
            // within 1e-8, as outer2 does)
            for (int k = 0; k < 3; ++k) u[k] /= rij;
            if (std::fabs(u[2]) > 0.99999999) u[2] = std::copysign(1.0, u[2]);
            for (int k = 0; k < 3 && li > 1; ++k) e1b[pidx(1 + k, 0)] = -B.tore * u[k] * sp_i;
            for (int k = 0; k < 3 && lj > 1; ++k) e2a[pidx(1 + k, 0)] = -A.tore * u[k] * sp_j;
          }
          for (int a = 0; a < li; ++a) e1b[pidx(a, a)] = -w1 * B.tore;
          for (int a = 0; a < lj; ++a) e2a[pidx(a, a)] = -w1 * A.tore;
          enuc = A.tore * B.tore * w1;
        }
        rowsum += enuc;
        for (int q = 0; q < 10; ++q) e1b_acc[(size_t)i * 10 + q] += e1b[q];
        const long long pid = (long long)i * (i - 1) / 2 + j;
        for (int q = 0; q < 10; ++q) e2a_pair[(size_t)pid * 10 + q] = e2a[q];
      }
      enuc_row[i] = rowsum;
    }
  };
  unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (numat < 64) nthr = 1;
  std::vector<std::thread> pool;
  for (unsigned t2 = 1; t2 < nthr; ++t2) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  double en = 0.0;
  for (int i = 0; i < numat; ++i) {
    en += enuc_row[i];
    const int li = g_elem[nat[i]].natorb;
    for (int a = 0; a < li; ++a)
      for (int b = 0; b <= a; ++b) h[nij(i, i) + pidx(a, b)] += e1b_acc[(size_t)i * 10 + pidx(a, b)];
  }
  for (int i = 0; i < numat; ++i)
    for (int j = 0; j < i; ++j) {
      const long long pid = (long long)i * (i - 1) / 2 + j;
      const int lj = g_elem[nat[j]].natorb;
      for (int a = 0; a < lj; ++a)
        for (int b = 0; b <= a; ++b) h[nij(j, j) + pidx(a, b)] += e2a_pair[(size_t)pid * 10 + pidx(a, b)];
    }
  *enuclr = en;
  return 0;
}

// The atom-level part of mhip_pm7_hamiltonian: orbital ranges, the
// isolated-atom reference heat and the diagonal starting density, with no
// pair work (for systems whose pairs are built sparsely).
int mhip_pm7_atoms(int numat, const int* nat, int charge, int* nfirst, int* nlast, double* atheat,
                   double* pdiag) {
  init_params();
  int norbs = 0;
  long long nw = 0;
  if (mhip_pm7_sizes(numat, nat, &norbs, &nw) != 0) return -1;
  int k = 1;
  double eheat = 0.0, eisol = 0.0;
  for (int i = 0; i < numat; ++i) {
    const Element& E = g_elem[nat[i]];
    nfirst[i] = k;
    k += E.natorb;
    nlast[i] = k - 1;
    eheat += E.p.dhf_atom;
    eisol += E.eisol;
  }
  *atheat = eheat - eisol * kFPC9;
  const double yy = (double)charge / (norbs + 1.0e-10);
  for (int i = 0; i < numat; ++i) {
    const Element& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a)
      pdiag[nfirst[i] - 1 + a] = (E.natorb == 1) ? E.tore - yy : E.tore * 0.25 - yy;
  }
  return 0;
}

// Two-electron integrals of the stored pairs only, in hcore order (per atom
// i: the blocks of its stored pairs j < i, then its one-centre block), as
// hcore_for_MOZYME keeps them for SEMIDIRECT.  w == nullptr: size only.
int mhip_pm7_integrals_mozyme(int numat, const int* nat, const double* coord, const int* nijbo,
                              long long* nw, double* w) {
  init_params();
  int norbs = 0;
  long long nfull = 0;
  if (!nw || mhip_pm7_sizes(numat, nat, &norbs, &nfull) != 0) return -1;
  auto nij = [&](int a, int b) { return nijbo[a + (size_t)b * numat]; };
  std::vector<long long> off(numat + 1);
  long long kr = 0;
  for (int i = 0; i < numat; ++i) {
    off[i] = kr;
    const int li = g_elem[nat[i]].natorb * (g_elem[nat[i]].natorb + 1) / 2;
    for (int j = 0; j < i; ++j)
      if (nij(i, j) >= 0) kr += (long long)li * (g_elem[nat[j]].natorb * (g_elem[nat[j]].natorb + 1) / 2);
    kr += (long long)li * li;
  }
  off[numat] = kr;
  if (!w) {
    *nw = kr;
    return 0;
  }
  if (*nw < kr) return -1;
  std::atomic<int> next_row{0};
  auto worker = [&]() {
    for (int i = next_row++; i < numat; i = next_row++) {
      const Element& A = g_elem[nat[i]];
      const int limij = A.natorb * (A.natorb + 1) / 2;
      long long kk = off[i];
      for (int j = 0; j < i; ++j) {
        if (nij(i, j) < 0) continue;
        const Element& B = g_elem[nat[j]];
        const int limkl = B.natorb * (B.natorb + 1) / 2;
        PairOut po;
        pair_integrals(A, B, coord + 3 * i, coord + 3 * j, po);
        for (int q = 0; q < limij * limkl; ++q) w[kk + q] = po.w[q];
        kk += (long long)limij * limkl;
      }
      one_center_block(A, w + kk);
    }
  };
  unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (numat < 64) nthr = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nthr; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  *nw = kr;
  return 0;
}

int mhip_pm7_pair(int zi, int zj, const double* xi, const double* xj, double* w, double* e1b, double* e2a,
                  double* enuc, double* di) {
  init_params();
  if (zi < 0 || zi >= 9 || zj < 0 || zj >= 9 || !g_elem[zi].ok || !g_elem[zj].ok) return -1;
  if (!xi || !xj || !w || !e1b || !e2a || !enuc || !di) return -1;
  const Element& A = g_elem[zi];
  const Element& B = g_elem[zj];
  PairOut po;
  pair_integrals(A, B, xi, xj, po);
  const int limij = A.natorb * (A.natorb + 1) / 2, limkl = B.natorb * (B.natorb + 1) / 2;
  std::memcpy(w, po.w, sizeof(double) * limij * limkl);
  std::memcpy(e1b, po.e1b, sizeof(double) * 10);
  std::memcpy(e2a, po.e2a, sizeof(double) * 10);
  *enuc = po.enuc;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) di[a * 4 + b] = po.di[a][b];
  return 0;
}

int mhip_pm7_element(int z, double* out) {
  init_params();
  if (z < 0 || z >= 9 || !g_elem[z].ok || !out) return -1;
  const Element& E = g_elem[z];
  out[0] = E.natorb;
  out[1] = E.tore;
  out[2] = E.am;
  out[3] = E.ad;
  out[4] = E.d1;
  out[5] = E.aq;
  out[6] = E.d2;
  out[7] = E.eisol;
  out[8] = E.p.dhf_atom;
  return 0;
}

}  // extern "C"
