// PM7 NDDO Hamiltonian for the s/p main-group elements H, C, N, O.
//
// This is the producer of the SCF inner loop's inputs: the one-electron matrix
// H (packed lower triangle), the two-electron integrals W in the order the
// Fock builder consumes them, the core-core repulsion and the isolated-atom
// reference energy.  It runs once per geometry on the host, exactly where the
// reference runs it (hcore before iter), and is not on the per-iteration path.
//
// Reference behaviour followed (all paths relative to the reference root):
//   hcore         src/integrals/hcore.F90:157-288     (H assembly, W order)
//   rotate        src/integrals/rotate.F90:16-124     (pair driver)
//   rotatd        src/integrals/mndod.F90:1096-1372   (local->molecular W, core)
//   reppd         src/integrals/mndod.F90:426-845     (22 local sp integrals)
//   spcore        src/integrals/mndod.F90:1683-1776   (electron-core integrals)
//   elenuc        src/integrals/mndod.F90:2367-2490   (core terms to H blocks)
//   rotmat        src/integrals/mndod.F90:1382-1569   (bond frame)
//   to_point      src/integrals/mndod.F90:3462-3486   (PM7 point-charge feathering)
//   ccrep         src/integrals/ccrep.F90:16-230      (PM7 core-core)
//   h1elec/diat   src/integrals/h1elec.F90, diat.F90, set.F90, coe.F90 (overlaps)
//   wstore        src/integrals/mndod.F90:2027-2095   (one-centre block)
//   calpar/inid   src/models/calpar.F90, mndod.F90:16-61 (derived multipoles)
//   pdiag         src/moldat.F90:731-790              (diagonal starting density)
//   parameters    src/models/parameters_for_PM7_C.F90 (H:24-33, C/N/O:~80-140,
//                 alpb/xfac 1405-1483, v_par7 1392-1395); eheat parameters_C.F90:224-231
//   constants     src/conref_C.F90:27-45 (2018 CODATA set, the default)
//
// The integral formulas and parameter values must equal the reference's; the
// code around them (data layout, the direct 4-index rotation, the parallel pair
// loop) is this library's own.
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
constexpr double kTrunc1 = 7.0;               // PM7 point-charge switch radius (A)
constexpr double kTrunc2 = 0.22;              // and its Gaussian exponent

struct Elem {
  int z = 0, natorb = 0, npq = 0, ios = 0, iop = 0;
  double tore = 0;
  double uss = 0, upp = 0, betas = 0, betap = 0, zs = 0, zp = 0;
  double gss = 0, gsp = 0, gpp = 0, gp2 = 0, hsp = 0;
  double g1 = 0, g2 = 0, g3 = 0;  // first core-core Gaussian
  double eheat = 0;
  // derived (calpar / inid)
  double eisol = 0, dd = 0, qq = 0, am = 0, ad = 0, aq = 0;
  double po1 = 0, po2 = 0, po3 = 0, po7 = 0, po9 = 0, ddp2 = 0, ddp3 = 0;
  bool ok = false;
};

Elem g_elem[9];
bool g_init = false;

// PM7 diatomic core-core pair parameters (alpha_AB, x_AB), symmetric.
double g_alpb[9][9], g_xfac[9][9];
// PM7 global core-core parameters v_par7(1..4).
constexpr double kPar1 = 8.947612, kPar2 = 6.024265, kPar3 = -0.012037, kPar4 = 0.701333;

void derive(Elem& e) {
  const int i = e.z;
  if (i == 1) {
    e.eisol = e.uss;
    e.am = e.gss / kEV;
    e.ad = e.am;
    e.aq = e.am;
  } else {
    const double gssc = std::max(e.ios - 1, 0);
    const int k = e.iop;
    const double gspc = e.ios * k;
    const int l = std::min(k, 6 - k);
    const double gp2c = (k * (k - 1)) / 2 + 0.5 * (l * (l - 1)) / 2;
    const double gppc = -0.5 * (l * (l - 1)) / 2;
    const double hspc = -k * e.ios * 0.5;
    e.zp = std::max(0.3, e.zp);
    double hpp = 0.5 * (e.gpp - e.gp2);
    hpp = std::max(0.1, hpp);
    e.hsp = std::max(1.0e-7, e.hsp);
    e.eisol = e.uss * e.ios + e.upp * e.iop + e.gss * gssc + e.gpp * gppc + e.gsp * gspc +
              e.gp2 * gp2c + e.hsp * hspc;
    const double qn = 2.0;  // principal quantum number of the 2sp shell
    e.dd = (2.0 * qn + 1) * std::pow(4.0 * e.zs * e.zp, qn + 0.5) /
           std::pow(e.zs + e.zp, 2.0 * qn + 2) / std::sqrt(3.0);
    e.qq = std::sqrt((4.0 * qn * qn + 6.0 * qn + 2.0) / 20.0) / e.zp;
    // Secant solve for the dipole additive term (calpar.F90, jmax = 5).
    double gdd1 = std::pow(e.hsp / (kEV * e.dd * e.dd), 1.0 / 3.0);
    double d1 = gdd1, d2 = gdd1 + 0.04;
    for (int j = 0; j < 5; ++j) {
      const double df = d2 - d1;
      const double h1 = 0.5 * d1 - 0.5 / std::sqrt(4.0 * e.dd * e.dd + 1.0 / (d1 * d1));
      const double h2 = 0.5 * d2 - 0.5 / std::sqrt(4.0 * e.dd * e.dd + 1.0 / (d2 * d2));
      if (std::fabs(h2 - h1) < 1.0e-25) break;
      const double d3 = d1 + df * (e.hsp / kEV - h1) / (h2 - h1);
      d1 = d2;
      d2 = d3;
    }
    // ... and for the quadrupole additive term.
    const double p4 = 16.0;
    double gqq = std::pow(p4 * hpp / (kEV * 48.0 * std::pow(e.qq, 4)), 0.2);
    double q1 = gqq, q2 = gqq + 0.04;
    for (int j = 0; j < 5; ++j) {
      const double qf = q2 - q1;
      const double h1 = 0.25 * q1 - 0.5 / std::sqrt(4.0 * e.qq * e.qq + 1.0 / (q1 * q1)) +
                        0.25 / std::sqrt(8.0 * e.qq * e.qq + 1.0 / (q1 * q1));
      const double h2 = 0.25 * q2 - 0.5 / std::sqrt(4.0 * e.qq * e.qq + 1.0 / (q2 * q2)) +
                        0.25 / std::sqrt(8.0 * e.qq * e.qq + 1.0 / (q2 * q2));
      if (std::fabs(h2 - h1) < 1.0e-25) break;
      const double q3 = q1 + qf * (hpp / kEV - h1) / (h2 - h1);
      q1 = q2;
      q2 = q3;
    }
    e.am = e.gss / kEV;
    e.ad = d2;
    e.aq = q2;
  }
  // inid: additive terms as radii, dipole / quadrupole separations
  e.po1 = 0.5 / e.am;
  if (e.ad > 1e-5) e.po2 = 0.5 / e.ad;
  if (e.aq > 1e-5) e.po3 = 0.5 / e.aq;
  e.po7 = e.po1;
  e.ddp2 = e.dd;
  e.ddp3 = e.qq * std::sqrt(2.0);
  e.po9 = e.po1;  // no separate core additive term for these elements in PM7
  if (i == 1) {
    e.po2 = 0.0;
    e.po3 = 0.0;
  }
  e.ok = true;
}

void init_params() {
  if (g_init) return;
  auto set = [](int z, int natorb, int npq, double tore, int ios, int iop, double uss, double upp,
                double betas, double betap, double zs, double zp, double gss, double gsp,
                double gpp, double gp2, double hsp, double g1, double g2, double g3,
                double eheat) {
    Elem& e = g_elem[z];
    e.z = z; e.natorb = natorb; e.npq = npq; e.tore = tore; e.ios = ios; e.iop = iop;
    e.uss = uss; e.upp = upp; e.betas = betas; e.betap = betap; e.zs = zs; e.zp = zp;
    e.gss = gss; e.gsp = gsp; e.gpp = gpp; e.gp2 = gp2; e.hsp = hsp;
    e.g1 = g1; e.g2 = g2; e.g3 = g3; e.eheat = eheat;
    derive(e);
  };
  //        Z  nao npq tore s  p     Uss          Upp          beta_s       beta_p       zeta_s     zeta_p
  //                               Gss         Gsp         Gpp         Gp2         Hsp
  //                               Gauss(1): K, L, M                    dHf(atom)
  set(1, 1, 1, 1.0, 1, 0, -11.070112, 0.0, -8.389745, 0.0, 1.260237, 0.0,
      14.149656, 0.0, 0.0, 0.0, 0.0, 0.177854, 1.428710, 0.991324, 52.102);
  set(6, 4, 2, 4.0, 2, 2, -51.372620, -40.135421, -14.414930, -7.893717, 1.942244, 1.708723,
      12.347323, 11.932801, 10.452226, 9.385498, 0.802634, 0.045888, 5.037055, 1.588715, 170.890);
  set(7, 4, 2, 5.0, 2, 3, -61.623260, -48.949816, -22.110958, -15.464433, 2.354344, 2.028288,
      11.881002, 9.572117, 12.210140, 10.299947, 2.977433, 0.015143, 4.734148, 1.516714, 113.000);
  set(8, 4, 2, 6.0, 2, 4, -96.087736, -71.087738, -67.776339, -20.267981, 5.972309, 2.349017,
      14.955121, 16.088521, 12.403337, 10.499706, 5.028656, -0.016243, 1.871970, 1.844360, 59.559);
  std::memset(g_alpb, 0, sizeof(g_alpb));
  std::memset(g_xfac, 0, sizeof(g_xfac));
  auto pair = [](int a, int b, double alpb, double xfac) {
    g_alpb[a][b] = g_alpb[b][a] = alpb;
    g_xfac[a][b] = g_xfac[b][a] = xfac;
  };
  pair(1, 1, 4.051163, 2.845627);
  pair(6, 1, 1.038716, 0.204582);
  pair(6, 6, 2.655746, 0.937818);
  pair(7, 1, 1.049958, 0.173966);
  pair(7, 6, 2.734636, 0.959163);
  pair(7, 7, 2.786676, 0.863430);
  pair(8, 1, 1.508030, 0.160602);
  pair(8, 6, 2.850931, 0.826848);
  pair(8, 7, 2.841313, 0.699190);
  pair(8, 8, 2.457135, 0.356439);
  g_init = true;
}

inline bool supported(int z) { return z == 1 || z == 6 || z == 7 || z == 8; }

// PM7 smooth switch from NDDO to point-charge electrostatics: returns the NDDO
// weight; `point` is the point-charge e-e value in eV.
inline double feather(double r_ang, double& point) {
  point = kEV * kA0 / r_ang;
  if (r_ang < kTrunc1) return 1.0 - std::exp(-(r_ang - kTrunc1) * (r_ang - kTrunc1) * kTrunc2);
  return 0.0;
}

// 22 unique two-centre two-electron integrals in the bond frame (eV), local
// axis 1 along the bond, plus the core-core Coulomb function gab.
void local_repulsion(const Elem& A, const Elem& B, double rij, double ri[22], double& gab) {
  const double ev1 = kEV / 2, ev2 = ev1 / 2, ev3 = ev2 / 2, ev4 = ev3 / 2;
  const double td = 2.0, pp = 0.5;
  for (int k = 0; k < 22; ++k) ri[k] = 0.0;
  const double r = rij / kA0;
  const bool si = A.natorb >= 3, sj = B.natorb >= 3;
  double aee = A.po9 + B.po9;
  aee = aee * aee;
  gab = kEV / std::sqrt(r * r + aee);
  aee = pp / A.am + pp / B.am;
  aee = aee * aee;
  if (!si && !sj) {
    ri[0] = kEV / std::sqrt(r * r + aee);
  } else if (si && !sj) {
    const double da = A.dd, qa = A.qq * td;
    double ade = pp / A.ad + pp / B.am; ade *= ade;
    double aqe = pp / A.aq + pp / B.am; aqe *= aqe;
    const double rsq = r * r;
    double arg[7], sq[7], x;
    arg[0] = rsq + aee;
    x = r + da; arg[1] = x * x + ade;
    x = r - da; arg[2] = x * x + ade;
    x = r + qa; arg[3] = x * x + aqe;
    x = r - qa; arg[4] = x * x + aqe;
    arg[5] = rsq + aqe;
    arg[6] = arg[5] + qa * qa;
    for (int k = 0; k < 7; ++k) sq[k] = std::sqrt(arg[k]);
    const double ee = kEV / sq[0];
    ri[0] = ee;
    ri[1] = ev1 / sq[1] - ev1 / sq[2];
    ri[2] = ee + ev2 / sq[3] + ev2 / sq[4] - ev1 / sq[5];
    ri[3] = ee + ev1 / sq[6] - ev1 / sq[5];
  } else if (!si && sj) {
    const double db = B.dd, qb = B.qq * td;
    double aed = pp / A.am + pp / B.ad; aed *= aed;
    double aeq = pp / A.am + pp / B.aq; aeq *= aeq;
    const double rsq = r * r;
    double arg[7], sq[7], x;
    arg[0] = rsq + aee;
    x = r - db; arg[1] = x * x + aed;
    x = r + db; arg[2] = x * x + aed;
    x = r - qb; arg[3] = x * x + aeq;
    x = r + qb; arg[4] = x * x + aeq;
    arg[5] = rsq + aeq;
    arg[6] = arg[5] + qb * qb;
    for (int k = 0; k < 7; ++k) sq[k] = std::sqrt(arg[k]);
    const double ee = kEV / sq[0];
    ri[0] = ee;
    ri[4] = ev1 / sq[1] - ev1 / sq[2];
    ri[10] = ee + ev2 / sq[3] + ev2 / sq[4] - ev1 / sq[5];
    ri[11] = ee + ev1 / sq[6] - ev1 / sq[5];
  } else {
    const double da = A.dd, db = B.dd;
    double qa = A.qq * td, qb = B.qq * td;
    double ade = pp / A.ad + pp / B.am; ade *= ade;
    double aqe = pp / A.aq + pp / B.am; aqe *= aqe;
    double aed = pp / A.am + pp / B.ad; aed *= aed;
    double aeq = pp / A.am + pp / B.aq; aeq *= aeq;
    double axx = pp / A.ad + pp / B.ad; axx *= axx;
    double adq = pp / A.ad + pp / B.aq; adq *= adq;
    double aqd = pp / A.aq + pp / B.ad; aqd *= aqd;
    double aqq = pp / A.aq + pp / B.aq; aqq *= aqq;
    const double rsq = r * r;
    double a[73], s[73], x, y, z, w;
    a[1] = rsq + aee;
    x = r + da; a[2] = x * x + ade;
    x = r - da; a[3] = x * x + ade;
    x = r - qa; a[4] = x * x + aqe;
    x = r + qa; a[5] = x * x + aqe;
    a[6] = rsq + aqe;
    a[7] = a[6] + qa * qa;
    x = r - db; a[8] = x * x + aed;
    x = r + db; a[9] = x * x + aed;
    x = r - qb; a[10] = x * x + aeq;
    x = r + qb; a[11] = x * x + aeq;
    a[12] = rsq + aeq;
    a[13] = a[12] + qb * qb;
    x = da - db; a[14] = rsq + axx + x * x;
    x = da + db; a[15] = rsq + axx + x * x;
    x = r + da - db; a[16] = x * x + axx;
    x = r - da + db; a[17] = x * x + axx;
    x = r - da - db; a[18] = x * x + axx;
    x = r + da + db; a[19] = x * x + axx;
    x = r + da; a[20] = x * x + adq;
    a[21] = a[20] + qb * qb;
    x = r - da; a[22] = x * x + adq;
    a[23] = a[22] + qb * qb;
    x = r - db; a[24] = x * x + aqd;
    a[25] = a[24] + qa * qa;
    x = r + db; a[26] = x * x + aqd;
    a[27] = a[26] + qa * qa;
    x = r + da - qb; a[28] = x * x + adq;
    x = r - da - qb; a[29] = x * x + adq;
    x = r + da + qb; a[30] = x * x + adq;
    x = r - da + qb; a[31] = x * x + adq;
    x = r + qa - db; a[32] = x * x + aqd;
    x = r + qa + db; a[33] = x * x + aqd;
    x = r - qa - db; a[34] = x * x + aqd;
    x = r - qa + db; a[35] = x * x + aqd;
    a[36] = rsq + aqq;
    x = qa - qb; a[37] = a[36] + x * x;
    x = qa + qb; a[38] = a[36] + x * x;
    a[39] = a[36] + qa * qa;
    a[40] = a[36] + qb * qb;
    a[41] = a[39] + qb * qb;
    x = r - qb; a[42] = x * x + aqq;
    a[43] = a[42] + qa * qa;
    x = r + qb; a[44] = x * x + aqq;
    a[45] = a[44] + qa * qa;
    x = r + qa; a[46] = x * x + aqq;
    a[47] = a[46] + qb * qb;
    x = r - qa; a[48] = x * x + aqq;
    a[49] = a[48] + qb * qb;
    x = r + qa - qb; a[50] = x * x + aqq;
    x = r + qa + qb; a[51] = x * x + aqq;
    x = r - qa - qb; a[52] = x * x + aqq;
    x = r - qa + qb; a[53] = x * x + aqq;
    qa = A.qq;
    qb = B.qq;
    x = da - qb; x = x * x;
    y = r - qb; y = y * y;
    z = da + qb; z = z * z;
    w = r + qb; w = w * w;
    a[54] = x + y + adq;
    a[55] = x + w + adq;
    a[56] = z + y + adq;
    a[57] = z + w + adq;
    x = qa - db; x = x * x;
    y = qa + db; y = y * y;
    z = r + qa; z = z * z;
    w = r - qa; w = w * w;
    a[58] = z + x + aqd;
    a[59] = w + x + aqd;
    a[60] = z + y + aqd;
    a[61] = w + y + aqd;
    x = qa - qb; x = x * x;
    a[62] = a[36] + td * x;
    y = qa + qb; y = y * y;
    a[63] = a[36] + td * y;
    a[64] = a[36] + td * (qa * qa + qb * qb);
    z = r + qa - qb; z = z * z;
    a[65] = z + x + aqq;
    a[66] = z + y + aqq;
    z = r + qa + qb; z = z * z;
    a[67] = z + x + aqq;
    a[68] = z + y + aqq;
    z = r - qa - qb; z = z * z;
    a[69] = z + x + aqq;
    a[70] = z + y + aqq;
    z = r - qa + qb; z = z * z;
    a[71] = z + x + aqq;
    a[72] = z + y + aqq;
    for (int k = 1; k <= 72; ++k) s[k] = std::sqrt(a[k]);
    const double ee = kEV / s[1];
    const double dze = -ev1 / s[2] + ev1 / s[3];
    const double qzze = ev2 / s[4] + ev2 / s[5] - ev1 / s[6];
    const double qxxe = ev1 / s[7] - ev1 / s[6];
    const double edz = -ev1 / s[8] + ev1 / s[9];
    const double eqzz = ev2 / s[10] + ev2 / s[11] - ev1 / s[12];
    const double eqxx = ev1 / s[13] - ev1 / s[12];
    const double dxdx = ev1 / s[14] - ev1 / s[15];
    const double dzdz = ev2 / s[16] + ev2 / s[17] - ev2 / s[18] - ev2 / s[19];
    const double dzqxx = ev2 / s[20] - ev2 / s[21] - ev2 / s[22] + ev2 / s[23];
    const double qxxdz = ev2 / s[24] - ev2 / s[25] - ev2 / s[26] + ev2 / s[27];
    const double dzqzz = -ev3 / s[28] + ev3 / s[29] - ev3 / s[30] + ev3 / s[31] - ev2 / s[22] + ev2 / s[20];
    const double qzzdz = -ev3 / s[32] + ev3 / s[33] - ev3 / s[34] + ev3 / s[35] + ev2 / s[24] - ev2 / s[26];
    const double qxxqxx = ev3 / s[37] + ev3 / s[38] - ev2 / s[39] - ev2 / s[40] + ev2 / s[36];
    const double qxxqyy = ev2 / s[41] - ev2 / s[39] - ev2 / s[40] + ev2 / s[36];
    const double qxxqzz = ev3 / s[43] + ev3 / s[45] - ev3 / s[42] - ev3 / s[44] - ev2 / s[39] + ev2 / s[36];
    const double qzzqxx = ev3 / s[47] + ev3 / s[49] - ev3 / s[46] - ev3 / s[48] - ev2 / s[40] + ev2 / s[36];
    const double qzzqzz = ev4 / s[50] + ev4 / s[51] + ev4 / s[52] + ev4 / s[53] - ev3 / s[48] -
                          ev3 / s[46] - ev3 / s[42] - ev3 / s[44] + ev2 / s[36];
    const double dxqxz = -ev2 / s[54] + ev2 / s[55] + ev2 / s[56] - ev2 / s[57];
    const double qxzdx = -ev2 / s[58] + ev2 / s[59] + ev2 / s[60] - ev2 / s[61];
    const double qxzqxz = ev3 / s[65] - ev3 / s[67] - ev3 / s[69] + ev3 / s[71] - ev3 / s[66] +
                          ev3 / s[68] + ev3 / s[70] - ev3 / s[72];
    ri[0] = ee;
    ri[1] = -dze;
    ri[2] = ee + qzze;
    ri[3] = ee + qxxe;
    ri[4] = -edz;
    ri[5] = dzdz;
    ri[6] = dxdx;
    ri[7] = -edz - qzzdz;
    ri[8] = -edz - qxxdz;
    ri[9] = -qxzdx;
    ri[10] = ee + eqzz;
    ri[11] = ee + eqxx;
    ri[12] = -dze - dzqzz;
    ri[13] = -dze - dzqxx;
    ri[14] = -dxqxz;
    ri[15] = ee + eqzz + qzze + qzzqzz;
    ri[16] = ee + eqzz + qxxe + qxxqzz;
    ri[17] = ee + eqxx + qzze + qzzqxx;
    ri[18] = ee + eqxx + qxxe + qxxqxx;
    ri[19] = qxzqxz;
    ri[20] = ee + eqxx + qxxe + qxxqyy;
    ri[21] = pp * (qxxqxx - qxxqyy);
  }
  // PM7 feathering toward point charges (monopole-like terms blend, others fade)
  double point;
  const double c = feather(rij, point);
  static const bool monopole[22] = {1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 1,
                                    1, 0, 0, 0, 1, 1, 1, 1, 0, 1, 0};
  for (int k = 0; k < 22; ++k) ri[k] = monopole[k] ? ri[k] * c + (1.0 - c) * point : ri[k] * c;
  gab = gab * c + (1.0 - c) * point;
}

// Electron-core attraction integrals in the bond frame: core[k][0] for the
// electrons of A in the field of core B (ss, s-sigma, sigma-sigma, pi-pi),
// core[k][1] for the electrons of B in the field of core A.  r in bohr.
void local_core(const Elem& A, const Elem& B, double r, double core[4][2]) {
  static const double pxy[7] = {1.0, -0.5, -0.5, 0.5, 0.25, 0.25, 0.5};
  for (int k = 0; k < 4; ++k) core[k][0] = core[k][1] = 0.0;
  const double r2 = r * r;
  const double aci = A.po9, acj = B.po9;
  const double ssi = (aci + B.po1) * (aci + B.po1);
  const double ssj = (acj + A.po1) * (acj + A.po1);
  core[0][0] = -B.tore * kEV / std::sqrt(r2 + ssj);
  core[0][1] = -A.tore * kEV / std::sqrt(r2 + ssi);
  if (A.z >= 3 || B.z >= 3) {
    if (A.z >= 3) {
      const double ppj = (acj + A.po7) * (acj + A.po7);
      const double da = A.ddp2, qa = A.ddp3 / std::sqrt(2.0), twoqa = qa + qa;
      const double adj = (A.po2 + acj) * (A.po2 + acj);
      const double aqj = (A.po3 + acj) * (A.po3 + acj);
      double x[7];
      x[0] = r2 + ppj;
      x[1] = r2 + aqj;
      x[2] = (r + da) * (r + da) + adj;
      x[3] = (r - da) * (r - da) + adj;
      x[4] = (r - twoqa) * (r - twoqa) + aqj;
      x[5] = (r + twoqa) * (r + twoqa) + aqj;
      x[6] = r2 + twoqa * twoqa + aqj;
      for (int k = 0; k < 7; ++k) x[k] = pxy[k] / std::sqrt(x[k]);
      core[1][0] = -B.tore * (x[2] + x[3]) * kEV;
      core[2][0] = -B.tore * (x[0] + x[1] + x[4] + x[5]) * kEV;
      core[3][0] = -B.tore * (x[0] + x[1] + x[6]) * kEV;
    }
    if (B.z >= 3) {
      const double ppi = (aci + B.po7) * (aci + B.po7);
      const double db = B.ddp2, qb = B.ddp3 / std::sqrt(2.0), twoqb = qb + qb;
      const double adi = (B.po2 + aci) * (B.po2 + aci);
      const double aqi = (B.po3 + aci) * (B.po3 + aci);
      double x[7];
      x[0] = r2 + ppi;
      x[1] = r2 + aqi;
      x[2] = (r + db) * (r + db) + adi;
      x[3] = (r - db) * (r - db) + adi;
      x[4] = (r - twoqb) * (r - twoqb) + aqi;
      x[5] = (r + twoqb) * (r + twoqb) + aqi;
      x[6] = r2 + twoqb * twoqb + aqi;
      for (int k = 0; k < 7; ++k) x[k] = pxy[k] / std::sqrt(x[k]);
      core[1][1] = -A.tore * (-(x[2] + x[3]) * kEV);
      core[2][1] = -A.tore * (x[0] + x[1] + x[4] + x[5]) * kEV;
      core[3][1] = -A.tore * (x[0] + x[1] + x[6]) * kEV;
    }
  }
}

// PM7 core-core repulsion (eV) for a pair at r bohr, given the feathered gab.
double core_core(const Elem& A, const Elem& B, double r_bohr, double gab) {
  const double r = r_bohr * kA0;
  const int ni = A.z, nj = B.z;
  const double enuc = A.tore * B.tore * gab;
  double fff = g_xfac[ni][nj];
  double abond = g_alpb[ni][nj];
  if (abond < 1e-6) abond = 1.2;
  double scale = 1.0 + 2.0 * fff * std::exp(-abond * (r + 0.0003 * std::pow(r, 6)));
  const int hi = std::max(ni, nj), lo = std::min(ni, nj);
  if (lo == 1) {
    if (hi == 6 || hi == 7) scale = 1.0 + 2.0 * fff * std::exp(-abond * r * r);
    if (hi == 8)
      scale = 1.0 + 2.0 * fff * std::exp(-abond * r * r) - kPar3 * std::exp(-kPar4 * r * 2);
  } else if (lo == 6 && hi == 6) {
    scale = scale + kPar1 * std::exp(-kPar2 * r);
  }
  double e = enuc * scale;
  double g = 0.0;
  double ax = A.g2 * (r - A.g3) * (r - A.g3);
  if (ax < 25.0) g += A.tore * B.tore / r * A.g1 * std::exp(-ax);
  ax = B.g2 * (r - B.g3) * (r - B.g3);
  if (ax < 25.0) g += A.tore * B.tore / r * B.g1 * std::exp(-ax);
  e += g;
  ax = r / (std::pow((double)ni, 0.3333) + std::pow((double)nj, 0.3333));
  if (ax < 3.0) e += std::min(1.0e-8 / std::pow(ax, 12), 1.0e5);
  return e;
}

// ---- Slater overlap integrals over s/p shells (first-row elements) ----------
constexpr double kFact[18] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0,
                              362880.0, 3628800.0, 39916800.0, 479001600.0, 6227020800.0,
                              87178291200.0, 1307674368000.0, 20922789888000.0,
                              355687428096000.0};

struct OvState {
  int isp = 1, ips = 2;
  double sa = 0, sb = 0;
  double a[8], b[8];
};

void a_ints(double x, int k, double* a) {
  const double c = std::exp(-x);
  a[0] = c / x;
  for (int i = 1; i <= k; ++i) a[i] = (a[i - 1] * i + c) / x;
}

void b_ints(double x, int k, double* b) {
  const double absx = std::fabs(x);
  int last = 0;
  bool expl = false, series = false;
  if (absx > 3.0) expl = true;
  else if (absx > 2.0) { if (k <= 10) expl = true; else { last = 15; series = true; } }
  else if (absx > 1.0) { if (k <= 7) expl = true; else { last = 12; series = true; } }
  else if (absx > 0.5) { if (k <= 5) expl = true; else { last = 7; series = true; } }
  else if (absx <= 1e-6) { }
  else { last = 6; series = true; }
  if (expl) {
    const double ex = std::exp(x), emx = 1.0 / ex;
    b[0] = (ex - emx) / x;
    for (int i = 1; i <= k; ++i) b[i] = (i * b[i - 1] + ((i % 2) ? -1.0 : 1.0) * ex - emx) / x;
  } else if (series) {
    for (int i = 0; i <= k; ++i) {
      double y = 0.0;
      for (int m = 0; m <= last; ++m) {
        const double xf = (m != 0) ? kFact[m] : 1.0;
        double pw = 1.0;
        for (int t = 0; t < m; ++t) pw *= -x;
        y += pw * (2 * ((m + i + 1) % 2)) / (xf * (m + i + 1));
      }
      b[i] = y;
    }
  } else {
    for (int i = 0; i <= k; ++i) b[i] = (2 * ((i + 1) % 2)) / (i + 1.0);
  }
}

void ov_set(OvState& st, double s1, double s2, int na, int nb, double rab, int ii) {
  if (na <= nb) { st.isp = 1; st.ips = 2; st.sa = s1; st.sb = s2; }
  else { st.isp = 2; st.ips = 1; st.sa = s2; st.sb = s1; }
  int j = ii + 2;
  if (ii > 3) j -= 1;
  const double alpha = 0.5 * rab * (st.sa + st.sb);
  const double beta = 0.5 * rab * (st.sb - st.sa);
  a_ints(alpha, j - 1, st.a);
  b_ints(beta, j - 1, st.b);
}

// Overlaps in the bond frame s[La][Lb][M] (La, Lb: 0=s, 1=p; M: 0=sigma, 1=pi).
void local_overlap(const Elem& A, const Elem& B, double r_ang, double s[2][2][2]) {
  for (int i = 0; i < 8; ++i) (&s[0][0][0])[i] = 0.0;
  auto cls = [](int z) { return z == 1 ? 1 : (z == 6 ? 4 : (z == 7 ? 5 : 6)); };
  const int jmax = std::max(cls(A.z), cls(B.z)), jmin = std::min(cls(A.z), cls(B.z));
  const int ii = (jmax == 1) ? 1 : (jmin == 1 ? 2 : 4);
  const int na = A.z, nb = B.z;
  const double esa = A.zs, epa = A.zp, esb = B.zs, epb = B.zp;
  const double rab = r_ang / kA0;
  OvState st;
  const double* a = st.a;
  const double* b = st.b;
  auto S = [&](int i, int j, int m) -> double& { return s[i - 1][j - 1][m - 1]; };
  if (ii == 1) {
    ov_set(st, esa, esb, na, nb, rab, ii);
    S(1, 1, 1) = 0.25 * std::sqrt(std::pow(st.sa * st.sb * rab * rab, 3)) *
                 (a[2] * b[0] - b[2] * a[0]);
  } else if (ii == 2) {
    ov_set(st, esa, esb, na, nb, rab, ii);
    const double rab4 = std::pow(rab, 4) * 0.125;
    double w = std::sqrt(std::pow(st.sa, 3) * std::pow(st.sb, 5)) * rab4;
    S(1, 1, 1) = std::sqrt(1.0 / 3.0) * w * (a[3] * b[0] - b[3] * a[0] + a[2] * b[1] - b[2] * a[1]);
    if (na > 1) ov_set(st, epa, esb, na, nb, rab, ii);
    if (nb > 1) ov_set(st, esa, epb, na, nb, rab, ii);
    w = std::sqrt(std::pow(st.sa, 3) * std::pow(st.sb, 5)) * rab4;
    S(st.isp, st.ips, 1) = w * (a[2] * b[0] - b[2] * a[0] + a[3] * b[1] - b[3] * a[1]);
  } else {
    ov_set(st, esa, esb, na, nb, rab, ii);
    const double rab4 = std::pow(rab, 5) * 0.0625;
    double w = std::sqrt(std::pow(st.sa * st.sb, 5)) * rab4;
    S(1, 1, 1) = w * (a[4] * b[0] + b[4] * a[0] - 2.0 * a[2] * b[2]) / 3.0;
    ov_set(st, esa, epb, na, nb, rab, ii);
    if (na > nb) ov_set(st, epa, esb, na, nb, rab, ii);
    w = std::sqrt(std::pow(st.sa * st.sb, 5)) * rab4;
    const double rt3 = 1.0 / std::sqrt(3.0);
    double d = a[3] * (b[0] - b[2]) - a[1] * (b[2] - b[4]);
    double e = b[3] * (a[0] - a[2]) - b[1] * (a[2] - a[4]);
    S(st.isp, st.ips, 1) = w * rt3 * (d + e);
    ov_set(st, epa, esb, na, nb, rab, ii);
    if (na > nb) ov_set(st, esa, epb, na, nb, rab, ii);
    w = std::sqrt(std::pow(st.sa * st.sb, 5)) * rab4;
    d = a[3] * (b[0] - b[2]) - a[1] * (b[2] - b[4]);
    e = b[3] * (a[0] - a[2]) - b[1] * (a[2] - a[4]);
    S(st.ips, st.isp, 1) = w * rt3 * (d - e);
    ov_set(st, epa, epb, na, nb, rab, ii);
    w = std::sqrt(std::pow(st.sa * st.sb, 5)) * rab4;
    S(2, 2, 1) = -w * (b[2] * (a[4] + a[0]) - a[2] * (b[4] + b[0]));
    S(2, 2, 2) = 0.5 * w * (a[4] * (b[0] - b[2]) - b[4] * (a[0] - a[2]) - a[2] * b[0] + b[2] * a[0]);
  }
}

// Molecular-frame overlap block di[a][b] (a on atom A, b on atom B; s,px,py,pz)
// for the vector v = xB - xA (Angstrom).  Returns false when beyond the cutoff.
void molecular_overlap(const Elem& A, const Elem& B, const double v[3], double di[4][4]) {
  for (int i = 0; i < 16; ++i) (&di[0][0])[i] = 0.0;
  const double x2 = v[0], y2 = v[1], z2 = v[2];
  const double rsq = x2 * x2 + y2 * y2 + z2 * z2;
  if (rsq >= 225.0) return;  // overlap cutoff (15 A)
  double xy = x2 * x2 + y2 * y2;
  const double r = std::sqrt(xy + z2 * z2);
  xy = std::sqrt(xy);
  double ca, cb, sa, sb;
  if (xy >= 1e-10) {
    ca = x2 / xy; cb = z2 / r; sa = y2 / xy; sb = xy / r;
  } else if (z2 < 0.0) {
    ca = -1.0; cb = -1.0; sa = 0.0; sb = 0.0;
  } else if (z2 == 0.0) {
    ca = cb = sa = sb = 0.0;
  } else {
    ca = 1.0; cb = 1.0; sa = 0.0; sb = 0.0;
  }
  if (r < 0.001) return;
  double s[2][2][2];
  local_overlap(A, B, r, s);
  // Rotation coefficients for p: sigma (c3), pi (c4), pi' (c2), per molecular
  // orbital x, y, z (orbital slots 1, 2, 3).
  const double c3[3] = {ca * sb, sa * sb, cb};
  const double c4[3] = {ca * cb, sa * cb, -sb};
  const double c2[3] = {-sa, ca, 0.0};
  const int la = A.natorb > 1 ? 2 : 1, lb = B.natorb > 1 ? 2 : 1;
  for (int i = 0; i < la; ++i) {
    for (int j = 0; j < lb; ++j) {
      const double aa = (j == 1) ? -1.0 : 1.0;
      const double bb = 1.0;
      const int na_ = (i == 0) ? 1 : 3, nb_ = (j == 0) ? 1 : 3;
      for (int k = 0; k < na_; ++k) {
        for (int l = 0; l < nb_; ++l) {
          const double s3a = (i == 0) ? 1.0 : c3[k], s3b = (j == 0) ? 1.0 : c3[l];
          const double p4a = (i == 0) ? 0.0 : c4[k], p4b = (j == 0) ? 0.0 : c4[l];
          const double p2a = (i == 0) ? 0.0 : c2[k], p2b = (j == 0) ? 0.0 : c2[l];
          const int ia = (i == 0) ? 0 : 1 + k, jb = (j == 0) ? 0 : 1 + l;
          di[ia][jb] = s[i][j][0] * (s3a * s3b) * aa + s[i][j][1] * (p4a * p4b + p2a * p2b) * bb;
        }
      }
    }
  }
}

// Index of the unordered orbital pair (i >= j, 0-based, s,px,py,pz) in the
// packed 10-element pair list: (s,s) (px,s) (px,px) (py,s) (py,px) ...
inline int pidx(int i, int j) {
  if (i < j) std::swap(i, j);
  return i * (i + 1) / 2 + j;
}

struct PairOut {
  double w[100];
  double e1b[10], e2a[10];
  double di[4][4];
  double enuc;
};

// Everything one atom pair contributes: W block, both core-attraction blocks,
// the resonance block and the core-core term.  Atom A has the larger index.
void pair_integrals(const Elem& A, const Elem& B, const double* xa, const double* xb,
                    PairOut& out) {
  const int li = A.natorb, lj = B.natorb;
  const int limij = li * (li + 1) / 2, limkl = lj * (lj + 1) / 2;
  std::memset(&out, 0, sizeof(out));
  const double d[3] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]};
  const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  // Resonance block (computed even when the two-electron part is skipped).
  if (rsq <= 1.0e4) {
    double di[4][4];
    molecular_overlap(A, B, d, di);
    const double bi[4] = {A.betas * 0.5, A.betap * 0.5, A.betap * 0.5, A.betap * 0.5};
    const double bj[4] = {B.betas * 0.5, B.betap * 0.5, B.betap * 0.5, B.betap * 0.5};
    for (int a = 0; a < li; ++a)
      for (int b = 0; b < lj; ++b) out.di[a][b] = di[a][b] * (bi[a] + bj[b]);
  }
  if (rsq < 0.00002) return;
  // Bond frame: row 0 along A->B, rows 1 and 2 perpendicular.
  double p[3][3];
  {
    const double x11 = d[0], x22 = d[1], x33 = d[2];
    const double bxy = x11 * x11 + x22 * x22;
    const double r = std::sqrt(bxy + x33 * x33);
    const double sqb = std::sqrt(bxy);
    double sb = sqb / r, ca, sa, cb;
    if (sb > 1.0e-7) {
      ca = x11 / sqb; sa = x22 / sqb; cb = x33 / r;
    } else {
      sa = 0.0; sb = 0.0;
      if (x33 < 0.0) { ca = -1.0; cb = -1.0; }
      else if (x33 > 0.0) { ca = 1.0; cb = 1.0; }
      else { ca = 0.0; cb = 0.0; }
    }
    p[0][0] = ca * sb; p[0][1] = sa * sb; p[0][2] = cb;
    p[1][0] = ca * cb; p[1][1] = sa * cb; p[1][2] = -sb;
    p[2][0] = -sa;     p[2][1] = ca;      p[2][2] = 0.0;
  }
  const double rij = std::sqrt(rsq);
  double ri[22], gab;
  local_repulsion(A, B, rij, ri, gab);
  const double rb = rij / kA0;
  double core[4][2];
  local_core(A, B, rb, core);
  {
    double point;
    const double c = feather(rij, point);
    const double pj = -(kEV / rb) * B.tore, pi = -(kEV / rb) * A.tore;
    core[0][0] = core[0][0] * c + (1.0 - c) * pj;
    core[1][0] = core[1][0] * c;
    core[2][0] = core[2][0] * c + (1.0 - c) * pj;
    core[3][0] = core[3][0] * c + (1.0 - c) * pj;
    core[0][1] = core[0][1] * c + (1.0 - c) * pi;
    core[1][1] = core[1][1] * c;
    core[2][1] = core[2][1] * c + (1.0 - c) * pi;
    core[3][1] = core[3][1] * c + (1.0 - c) * pi;
  }
  // Local (bond-frame) two-electron tensor over pairs: L[pairA][pairB], local
  // orbitals 0=s, 1=sigma, 2=pi, 3=pi'.
  double L[10][10];
  std::memset(L, 0, sizeof(L));
  auto setL = [&](int a, int b, int c, int dd, double v) { L[pidx(a, b)][pidx(c, dd)] = v; };
  setL(0, 0, 0, 0, ri[0]);
  setL(0, 0, 1, 0, ri[4]);  setL(0, 0, 1, 1, ri[10]);
  setL(0, 0, 2, 2, ri[11]); setL(0, 0, 3, 3, ri[11]);
  setL(1, 0, 0, 0, ri[1]);  setL(1, 0, 1, 0, ri[5]);  setL(1, 0, 1, 1, ri[12]);
  setL(1, 0, 2, 2, ri[13]); setL(1, 0, 3, 3, ri[13]);
  setL(1, 1, 0, 0, ri[2]);  setL(1, 1, 1, 0, ri[7]);  setL(1, 1, 1, 1, ri[15]);
  setL(1, 1, 2, 2, ri[17]); setL(1, 1, 3, 3, ri[17]);
  setL(2, 0, 2, 0, ri[6]);  setL(2, 0, 2, 1, ri[14]); setL(2, 1, 2, 0, ri[9]);
  setL(2, 1, 2, 1, ri[19]);
  setL(2, 2, 0, 0, ri[3]);  setL(2, 2, 1, 0, ri[8]);  setL(2, 2, 1, 1, ri[16]);
  setL(2, 2, 2, 2, ri[18]); setL(2, 2, 3, 3, ri[20]);
  setL(3, 0, 3, 0, ri[6]);  setL(3, 0, 3, 1, ri[14]); setL(3, 1, 3, 0, ri[9]);
  setL(3, 1, 3, 1, ri[19]); setL(3, 2, 3, 2, ri[21]);
  setL(3, 3, 0, 0, ri[3]);  setL(3, 3, 1, 0, ri[8]);  setL(3, 3, 1, 1, ri[16]);
  setL(3, 3, 2, 2, ri[20]); setL(3, 3, 3, 3, ri[18]);
  // Molecular orbital mu = sum_a U[mu][a] * local a.  The multipole
  // integrals above place the sigma axis from B toward A (the dipole charge
  // at r+D faces away from B), so the sigma column enters with the opposite
  // sign to the A->B frame used for the core and overlap terms.  This
  // convention is pinned by the reference's CO2 eigenvalues and H2O charges.
  double U[4][4] = {{1, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k < 3; ++k) U[1 + m][1 + k] = (k == 0 ? -1.0 : 1.0) * p[k][m];
  // Pair transform T[molpair][localpair] over unordered pairs (with the
  // off-diagonal local pair appearing twice).
  double TA[10][10], TB[10][10];
  auto build_T = [&](int norb, double T[10][10]) {
    std::memset(T, 0, sizeof(double) * 100);
    for (int mu = 0; mu < norb; ++mu)
      for (int nu = 0; nu <= mu; ++nu)
        for (int a = 0; a < norb; ++a)
          for (int b = 0; b < norb; ++b) T[pidx(mu, nu)][pidx(a, b)] += U[mu][a] * U[nu][b];
  };
  build_T(li, TA);
  build_T(lj, TB);
  // (mu nu | lam sig) = sum TA[mn][ab] L[ab][cd] TB[ls][cd]
  double M[10][10];
  for (int mn = 0; mn < limij; ++mn)
    for (int cd = 0; cd < limkl; ++cd) {
      double acc = 0.0;
      for (int ab = 0; ab < limij; ++ab) acc += TA[mn][ab] * L[ab][cd];
      M[mn][cd] = acc;
    }
  for (int mn = 0; mn < limij; ++mn)
    for (int ls = 0; ls < limkl; ++ls) {
      double acc = 0.0;
      for (int cd = 0; cd < limkl; ++cd) acc += M[mn][cd] * TB[ls][cd];
      out.w[mn * limkl + ls] = acc;
    }
  // Electron-core attraction blocks in the molecular frame.
  auto core_block = [&](int norb, int col, double* e) {
    for (int mu = 0; mu < norb; ++mu)
      for (int nu = 0; nu <= mu; ++nu) {
        double v;
        if (mu == 0) v = core[0][col];
        else if (nu == 0) v = p[0][mu - 1] * core[1][col];
        else
          v = core[2][col] * p[0][mu - 1] * p[0][nu - 1] +
              core[3][col] * (p[1][mu - 1] * p[1][nu - 1] + p[2][mu - 1] * p[2][nu - 1]);
        e[pidx(mu, nu)] = v;
      }
  };
  core_block(li, 0, out.e1b);
  core_block(lj, 1, out.e2a);
  out.enuc = core_core(A, B, rb, gab);
}

void one_center_block(const Elem& E, double* w) {
  const int ilim = E.natorb * (E.natorb + 1) / 2;
  for (int i = 0; i < ilim * ilim; ++i) w[i] = 0.0;
  auto W = [&](int i, int j) -> double& { return w[(j - 1) * ilim + (i - 1)]; };
  W(1, 1) = E.gss;
  if (E.natorb > 2) {
    const int ipx = 3, ipy = 6, ipz = 10;
    W(ipx, 1) = W(ipy, 1) = W(ipz, 1) = E.gsp;
    W(1, ipx) = W(1, ipy) = W(1, ipz) = E.gsp;
    W(ipx, ipx) = W(ipy, ipy) = W(ipz, ipz) = E.gpp;
    W(ipy, ipx) = W(ipz, ipx) = W(ipz, ipy) = E.gp2;
    W(ipx, ipy) = W(ipx, ipz) = W(ipy, ipz) = E.gp2;
    W(2, 2) = W(4, 4) = W(7, 7) = E.hsp;
    W(5, 5) = W(8, 8) = W(9, 9) = 0.5 * (E.gpp - E.gp2);
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
    const Elem& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a) {
      const long long o = nfirst[i] - 1 + a;
      h[o * (o + 1) / 2 + o] = (a == 0) ? E.uss : E.upp;
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
    const Elem& A = g_elem[nat[i]];
    const int li = A.natorb, limij = li * (li + 1) / 2;
    long long kk = off[i];
    double rowsum = 0.0;
    for (int j = 0; j < i; ++j) {
      const Elem& B = g_elem[nat[j]];
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
    eheat += g_elem[nat[i]].eheat;
    eisol += g_elem[nat[i]].eisol;
  }
  *atheat = eheat - eisol * kFPC9;
  // Diagonal starting density
  const double yy = (double)charge / (norbs + 1.0e-10);
  for (int i = 0; i < numat; ++i) {
    const Elem& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a)
      pdiag[nfirst[i] - 1 + a] = (E.natorb == 1) ? E.tore - yy : E.tore * 0.25 - yy;
  }
  return 0;
}

double mhip_ev_to_kcal(void) { return kFPC9; }

// The one-electron matrix on the MOZYME sparse storage and the core-core
// energy, as hcore_for_MOZYME forms them for a discrete system
// (src/MOZYME/hcore_for_MOZYME.F90:100-260): stored pairs (nijbo >= 0) take
// the NDDO terms (resonance block, electron-core attraction, core-core);
// pairs of the dipole shell (-2) outer2's point charge plus dipole
// (src/MOZYME/outer2.F90:16-150, from the NDDO (ss|ss), (s sigma|ss) and
// (ss|s sigma) integrals); the others outer1's point charge
// (src/MOZYME/outer1.F90:16-72) with PM7's switch to the bare Coulomb.
int mhip_pm7_hamiltonian_mozyme(int numat, const int* nat, const double* coord, const int* nijbo,
                                long long mpack, double* h, double* enuclr) {
  init_params();
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
    const Elem& E = g_elem[nat[i]];
    for (int a = 0; a < E.natorb; ++a) h[nij(i, i) + a * (a + 1) / 2 + a] = (a == 0) ? E.uss : E.upp;
  }
  const long long npairs = (long long)numat * (numat - 1) / 2;
  std::vector<double> e1b_acc((size_t)numat * 10, 0.0), enuc_row(numat, 0.0);
  std::vector<double> e2a_pair((size_t)npairs * 10, 0.0);
  std::atomic<int> next_row{0};
  auto worker = [&]() {
    for (int i = next_row++; i < numat; i = next_row++) {
      const Elem& A = g_elem[nat[i]];
      const int li = A.natorb;
      double rowsum = 0.0;
      for (int j = 0; j < i; ++j) {
        const Elem& B = g_elem[nat[j]];
        const int lj = B.natorb;
        double e1b[10] = {0}, e2a[10] = {0}, enuc = 0.0;
        const int o = nij(i, j);
        if (o >= 0) {
          PairOut po;
          pair_integrals(A, B, coord + 3 * i, coord + 3 * j, po);
          for (int a = 0; a < li; ++a)
            for (int b = 0; b < lj; ++b) h[o + a * lj + b] += po.di[a][b];
          std::memcpy(e1b, po.e1b, sizeof(e1b));
          std::memcpy(e2a, po.e2a, sizeof(e2a));
          enuc = po.enuc;
        } else {
          const double* xi = coord + 3 * i;
          const double* xj = coord + 3 * j;
          double x[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
          const double rij = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
          double w1;
          if (o == -2) {
            double ri[22], gab;
            local_repulsion(A, B, rij, ri, gab);
            w1 = ri[0];
            for (int k = 0; k < 3; ++k) x[k] /= rij;
            if (std::fabs(x[2]) > 0.99999999) x[2] = std::copysign(1.0, x[2]);
            // the (p, s) elements: outer2's -w tore with w = -ri(2) x (atom i) and
            // -ri(5) x (atom j); the sigma-odd integrals of local_repulsion carry
            // the opposite sign to the reference's (its sigma axis runs from B
            // to A), hence -ri here, which also makes these terms continue the
            // NDDO attraction across the edge of the stored pairs
            if (li > 1) {
              e1b[1] = -ri[1] * x[0] * B.tore;
              e1b[3] = -ri[1] * x[1] * B.tore;
              e1b[6] = -ri[1] * x[2] * B.tore;
            }
            if (lj > 1) {
              e2a[1] = -ri[4] * x[0] * A.tore;
              e2a[3] = -ri[4] * x[1] * A.tore;
              e2a[6] = -ri[4] * x[2] * A.tore;
            }
          } else {
            const double r = rij / kA0;
            double aee = 0.5 / A.am + 0.5 / B.am;
            aee = aee * aee;
            w1 = kEV / std::sqrt(r * r + aee);
            double point;
            const double cf = feather(rij, point);
            w1 = w1 * cf + (1.0 - cf) * point;
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
    const Elem& E = g_elem[nat[i]];
    nfirst[i] = k;
    k += E.natorb;
    nlast[i] = k - 1;
    eheat += E.eheat;
    eisol += E.eisol;
  }
  *atheat = eheat - eisol * kFPC9;
  const double yy = (double)charge / (norbs + 1.0e-10);
  for (int i = 0; i < numat; ++i) {
    const Elem& E = g_elem[nat[i]];
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
      const Elem& A = g_elem[nat[i]];
      const int limij = A.natorb * (A.natorb + 1) / 2;
      long long kk = off[i];
      for (int j = 0; j < i; ++j) {
        if (nij(i, j) < 0) continue;
        const Elem& B = g_elem[nat[j]];
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

int mhip_pm7_element(int z, double* out) {
  init_params();
  if (z < 0 || z >= 9 || !g_elem[z].ok || !out) return -1;
  const Elem& E = g_elem[z];
  out[0] = E.natorb;
  out[1] = E.tore;
  out[2] = E.am;
  out[3] = E.ad;
  out[4] = E.dd;
  return 0;
}

}  // extern "C"
