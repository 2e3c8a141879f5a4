// PM7's post-SCF corrections: functions of the geometry alone that the
// reference adds to the SCF heat of formation for PM7
// (src/corrections/post_scf_corrections.F90:112-114):
//   dispersion   a damped C6/R^6 pair sum (Grimme's 2004 form with Slater-
//                Kirkwood combined C6, PM7's damping and scale;
//                src/corrections/H_bond_correction_PM6_DH_Dispersion.F90:59-226);
//   hydrogen     Korth's third-generation DH+ term (J. Chem. Theory Comput.
//   bonds        6, 3808 (2010)) over every X-H...Y triple with X, Y in {N, O},
//                with PM7's damping and O-O short-range term
//                (src/corrections/Hydrogen_bond_corrections.F90:15-465,
//                H_bond_correction_EH_plus.F90:16-358,
//                H_bond_correction_bits.F90:16-264).
// Both run once per geometry on the host; neither touches the SCF.
//
// The reference's switches are discontinuous in places (which neighbours
// define an angle, which branch a torsion's sign selects, a 4e-5 rad floor
// under which a torsion counts as zero), so the selection rules, the order in
// which ties are broken and the torsion's construction follow the
// reference's steps; a handful of its angle constants are single-precision
// literals in the source and are used here at that precision.
#include <cmath>
#include <vector>
#include <algorithm>

#include "mopac_hip.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kA0 = 0.529177210903;
constexpr double kEV = 27.211386245988;
constexpr double kFPC9 = 23.060547830619029;

bool supported(int z) { return z == 1 || z == 6 || z == 7 || z == 8; }

struct Geometry {
  int n;
  const int* z;
  const double* x;  // 3 n, Angstrom
  double d2(int a, int b) const {
    const double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1], dz = x[3 * a + 2] - x[3 * b + 2];
    return dx * dx + dy * dy + dz * dz;
  }
  double dist(int a, int b) const { return std::sqrt(d2(a, b)); }
  // angle a-b-c at b from the three squared distances (law of cosines)
  double angle(int a, int b, int c) const {
    const double ab = d2(a, b), bc = d2(b, c), ac = d2(a, c);
    const double m = std::sqrt(ab * bc);
    if (m < 1.0e-20) return 0.0;
    const double t = std::max(-1.0, std::min(1.0, 0.5 * (ab + bc - ac) / m));
    return std::acos(t);
  }
  // Torsion a-b-c-d in [0, 2 pi): the b-c axis (about c) is turned onto z,
  // and the angle runs from d's projection to a's, 0 below 4e-5 rad or when a
  // projection vanishes.
  double torsion(int a, int b, int c, int d) const {
    double va[3], vb[3], vd[3];
    for (int k = 0; k < 3; ++k) {
      va[k] = x[3 * a + k] - x[3 * c + k];
      vb[k] = x[3 * b + k] - x[3 * c + k];
      vd[k] = x[3 * d + k] - x[3 * c + k];
    }
    const double len = std::sqrt(vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2]);
    const double cz = std::max(-1.0, std::min(1.0, vb[2] / len));
    const double s2 = 1.0 - cz * cz;
    double ax, ay, dx, dy, ct, st;
    const double rxy = (s2 > 0.0) ? len * std::sqrt(s2) : 0.0;
    if (rxy > 1.0e-6) {
      // turn about z so b lies in the yz plane, then about x onto z
      const double cph = vb[1] / rxy, sph = vb[0] / rxy;
      ax = va[0] * cph - va[1] * sph;
      dx = vd[0] * cph - vd[1] * sph;
      const double ay2 = va[0] * sph + va[1] * cph;
      const double by2 = vb[0] * sph + vb[1] * cph;
      const double dy2 = vd[0] * sph + vd[1] * cph;
      ct = cz;
      st = by2 / len;
      ay = ay2 * ct - va[2] * st;
      dy = dy2 * ct - vd[2] * st;
    } else {
      ct = cz;
      st = 0.0;
      ax = va[0];
      dx = vd[0];
      ay = va[1] * ct - va[2] * st;
      dy = vd[1] * ct - vd[2] * st;
    }
    // signed angle between (dx, dy) and (ax, ay)
    double t = 0.0;
    const double tiny = 1.0e-6;
    if ((std::fabs(dx) >= tiny || std::fabs(dy) >= tiny) && (std::fabs(ax) >= tiny || std::fabs(ay) >= tiny)) {
      const double nd = 1.0 / std::sqrt(dx * dx + dy * dy), na = 1.0 / std::sqrt(ax * ax + ay * ay);
      const double p1 = dx * nd, p2 = dy * nd, q1 = ax * na, q2 = ay * na;
      const double s = p1 * q2 - p2 * q1;
      const double c = std::max(-1.0, std::min(1.0, p1 * q1 + p2 * q2));
      const double r = std::acos(c);
      if (std::fabs(r) >= 4.0e-5) t = (s > 0.0) ? r : 2.0 * kPi - r;
    }
    if (t >= 6.28318530717959) t = 0.0;
    return t;
  }
};

// ---- bonding topology -------------------------------------------------------
// Atoms are bonded within a multiple of the sum of their covalent radii
// (src/corrections/set_up_dentate.F90:16-135, radii atomradii_C.F90:85-90):
// 1.25 for C-H, 1.2 for C-C and C-N, 1.1 otherwise; at most 15 bonds per
// atom; a hydrogen with two or more partners keeps only its non-hydrogen ones.
double covalent_radius(int z) {
  switch (z) {
    case 1: return 0.37;
    case 6: return 0.77;
    case 7: return 0.75;
    default: return 0.73;
  }
}
std::vector<int> bond_counts(const Geometry& g) {
  std::vector<std::vector<int>> nb(g.n);
  for (int i = 0; i < g.n; ++i)
    for (int j = 0; j < i; ++j) {
      const int lo = std::min(g.z[i], g.z[j]), hi = std::max(g.z[i], g.z[j]);
      double safety = 1.1;
      if (lo == 1 && hi == 6) safety = 1.25;
      if (lo == 6 && hi < 8) safety = 1.2;
      const double lim = safety * (covalent_radius(g.z[i]) + covalent_radius(g.z[j]));
      if (g.d2(i, j) < lim * lim && nb[i].size() < 15 && nb[j].size() < 15) {
        nb[i].push_back(j);
        nb[j].push_back(i);
      }
    }
  std::vector<int> count(g.n);
  for (int i = 0; i < g.n; ++i) {
    if (g.z[i] == 1 && nb[i].size() >= 2) {
      int k = 0;
      for (int j : nb[i]) k += (g.z[j] != 1);
      count[i] = k;
    } else {
      count[i] = (int)nb[i].size();
    }
  }
  return count;
}

// ---- dispersion ---------------------------------------------------------------
// E = -s6 sum_{i<j} C6_ij / R^6 f(R), f = 1 / (1 + exp(-alpha (R / (s R0) - 1))),
// C6_ij from the atoms' C6 and Slater-Kirkwood effective electron numbers,
// R0 from the atoms' van der Waals radii; kJ nm^6/mol -> kcal/mol.  A carbon
// with four bonds takes the saturated-carbon C6.
struct DispAtom {
  double c6, r0_pm, neff;
};
DispAtom disp_atom(int z, int nbonds) {
  switch (z) {
    case 1: return {0.16, 156.0, 0.80};
    case 6: return {nbonds == 4 ? 0.95 : 1.65, 170.0, 2.50};
    case 7: return {1.11, 155.0, 2.82};
    default: return {0.70, 152.0, 3.15};
  }
}
double dispersion(const Geometry& g, const std::vector<int>& nbonds) {
  const double alpha = 15.450118, s = 1.226593, s6 = 2.286419;
  double e = 0.0;
  for (int i = 0; i < g.n; ++i)
    for (int j = i + 1; j < g.n; ++j) {
      const double r2 = g.d2(i, j);
      if (!(r2 < 100.0 * 100.0)) continue;
      const DispAtom a = disp_atom(g.z[i], nbonds[i]), b = disp_atom(g.z[j], nbonds[j]);
      const double c6 = 2.0 * std::pow(a.c6 * a.c6 * b.c6 * b.c6 * a.neff * b.neff, 1.0 / 3.0) /
                        (std::pow(a.c6 * b.neff * b.neff, 1.0 / 3.0) + std::pow(b.c6 * a.neff * a.neff, 1.0 / 3.0));
      const double r0 = (a.r0_pm * a.r0_pm * a.r0_pm + b.r0_pm * b.r0_pm * b.r0_pm) /
                        (a.r0_pm * a.r0_pm + b.r0_pm * b.r0_pm) / 1000.0 * 2.0;  // nm
      const double r = std::sqrt(r2) * 0.1;   // nm
      const double f = 1.0 / (1.0 + std::exp(-alpha * (r / (s * r0) - 1.0)));
      e -= c6 / std::pow(r, 6) * f / (1000.0 * 4.184);
    }
  return e * s6;
}

// ---- hydrogen bonds (DH+) --------------------------------------------------------
// One candidate X-H...Y: X and Y are N or O, H lies within 1.4 A of X, Y
// within 7 A of X and the X-H-Y angle is above 90 degrees.  nx / ny: the
// neighbours that orient X and Y (the reference's hblist columns 2-4 / 6-8).
struct HBond {
  int x, h, y;
  int nx[3], ny[3];
  int nbx = 0, nby = 0;
  bool skip = false;  // X and Y bonded to each other or to a common atom
};

// Bond radii of the DH+ neighbour search: 4/3 of its own covalent table.
double dh_radius(int z) {
  switch (z) {
    case 1: return 4.0 / 3.0 * 0.32;
    case 6: return 4.0 / 3.0 * 0.75;
    case 7: return 4.0 / 3.0 * 0.71;
    default: return 4.0 / 3.0 * 0.63;
  }
}

std::vector<HBond> find_hbonds(const Geometry& g) {
  std::vector<int> acc, hyd;
  std::vector<char> used(g.n, 0);
  const double rah2 = 1.4 * 1.4, cut2 = 7.0 * 7.0;
  for (int i = 0; i < g.n; ++i) {
    if (g.z[i] != 7 && g.z[i] != 8) continue;
    acc.push_back(i);
    for (int j = 0; j < g.n; ++j)
      if (g.z[j] == 1 && !used[j] && g.d2(i, j) < rah2) {
        hyd.push_back(j);
        used[j] = 1;
      }
  }
  std::vector<HBond> out;
  for (int i : acc)
    for (int h : hyd) {
      if (!(g.d2(i, h) < rah2)) continue;
      for (int k : acc) {
        if (k == i || !(g.d2(k, i) < cut2)) continue;
        if (!(g.angle(k, h, i) > kPi * 0.5)) continue;
        bool dup = false;
        for (const HBond& b : out)
          if (b.h == h && ((b.x == k && b.y == i) || (b.x == i && b.y == k))) {
            dup = true;
            break;
          }
        if (dup) continue;
        HBond b;
        b.x = i;
        b.h = h;
        b.y = k;
        out.push_back(b);
      }
    }
  return out;
}

// Neighbour lists of X and Y and the three atoms that orient each
// (setup_DH_Plus, Hydrogen_bond_corrections.F90:159-465).
void orient(const Geometry& g, HBond& b) {
  auto bonded = [&](int a, int c) { return g.dist(a, c) < dh_radius(g.z[a]) + dh_radius(g.z[c]); };
  // atoms bonded to c (c itself excluded), in atom order; a fifth partner
  // drops the one farthest from c (the first of equals)
  auto neighbours = [&](int c, std::vector<int>& lst) {
    for (int j = 0; j < g.n; ++j) {
      if (j == c || !bonded(j, c)) continue;
      lst.push_back(j);
      if (lst.size() == 5) {
        int far = 0;
        double dmax = 0.0;
        for (int k = 0; k < 5; ++k)
          if (dmax < g.dist(c, lst[k])) dmax = g.dist(c, lst[k]), far = k;
        lst.erase(lst.begin() + far);
      }
    }
  };
  std::vector<int> lx, ly;
  neighbours(b.x, lx);
  neighbours(b.y, ly);
  b.nbx = (int)lx.size();
  b.nby = (int)ly.size();
  for (int a : lx) {
    if (a == b.y) b.skip = true;  // 1-3
    for (int c : ly)
      if (a == c && c != b.h) b.skip = true;  // 1-4
  }
  if (b.skip) return;
  // the neighbour of c farthest from H, excluding `not1` and `not2` (first of equals)
  auto farthest = [&](const std::vector<int>& lst, int not1, int not2) {
    int best = -1;
    double dmax = -1.0;
    for (int a : lst) {
      if (a == not1 || a == not2) continue;
      const double d = g.dist(a, b.h);
      if (d > dmax) dmax = d, best = a;
    }
    return best;
  };
  auto fill = [&](int c, const std::vector<int>& lst, int* o) {
    const int n = (int)lst.size();
    const int self_or_h = bonded(c, b.h) ? b.h : c;
    if (n >= 3) {
      o[0] = farthest(lst, -1, -1);
      o[1] = farthest(lst, o[0], -1);
      o[2] = farthest(lst, o[0], o[1]);
    } else if (n == 2) {
      o[0] = farthest(lst, -1, -1);
      o[1] = (lst[1] != o[0]) ? lst[1] : lst[0];
      o[2] = self_or_h;
    } else if (n == 1) {
      o[0] = lst[0];
      std::vector<int> l2;
      for (int k = 0; k < g.n; ++k)
        if (k != o[0] && bonded(k, o[0])) l2.push_back(k);
      o[1] = farthest(l2, -1, -1);
      o[2] = self_or_h;
    } else {
      o[0] = o[1] = o[2] = self_or_h;
    }
  };
  fill(b.x, lx, b.nx);
  fill(b.y, ly, b.ny);
}

inline double wrap_pi(double t) {
  if (t <= -kPi) t += 2.0 * kPi;
  if (t > kPi) t -= 2.0 * kPi;
  return t;
}
inline double deg(double a) { return kPi / (180.0 / a); }  // a degrees in radians, as the reference forms it

// Angular factors of one end (X with its neighbours, or Y): the cosine of the
// deviation of the n0-c-H angle from its target and of the n1-n0-c-H torsion
// from its target.  `x_side` selects the reference's X-end constants (its
// tetrahedral angles there are single-precision literals).
struct EndTerms {
  double angle_cos, torsion_cos;
  bool ok;
};
EndTerms end_terms(const Geometry& g, const HBond& b, int c, const int* o, int nbonds, bool x_side) {
  const double tet = x_side ? (double)109.48f : 109.48, tors = x_side ? (double)54.74f : 54.74;
  double a_shift = 0.0, a_shift2 = 0.0, t_shift = 0.0;
  bool nr3 = false, carbonyl = false;
  if (g.z[c] == 8) {
    if (nbonds == 1) {
      a_shift = kPi;
      a_shift2 = deg(120.0);
      carbonyl = true;
    } else {
      a_shift = a_shift2 = deg(tet);
      t_shift = deg(tors);
    }
  } else if (g.z[c] == 7) {
    if (nbonds == 2) {
      a_shift = a_shift2 = deg(120.0);
    } else {
      a_shift = a_shift2 = deg(tet);
      t_shift = deg(tors);
      nr3 = true;
    }
  }
  // an NR3 group interpolates its targets between tetrahedral and planar
  double tcheck = 0.0;
  if (nr3) {
    double t = wrap_pi(g.torsion(o[1], o[0], c, o[2]));
    t = (t < 0.0) ? -kPi - t : kPi - t;
    const double keep = t;
    const double td = std::fabs(t) * 180.0 / kPi;
    t_shift += deg((54.74 - td) / 54.74 * 35.26);
    a_shift -= deg((54.74 - td) / 54.74 * 19.48);
    a_shift2 = a_shift;
    tcheck = keep;
  }
  const double ang = g.angle(o[0], c, b.h);
  const double ac = std::max(std::cos(a_shift - ang), std::cos(a_shift2 - ang));
  if (ac <= 0.0) return {ac, 0.0, false};
  double tc = wrap_pi(g.torsion(o[1], o[0], c, b.h));
  if (!carbonyl || std::fabs(tc * 180.0 / kPi) > 90.0) tc = (tc < 0.0) ? -kPi - tc : kPi - tc;
  double tcos;
  if (tcheck < 0.0) {
    tcos = std::cos(wrap_pi(t_shift - tc));
  } else if (tcheck > 0.0) {
    tcos = std::cos(wrap_pi(-t_shift - tc));
  } else {
    tcos = std::max(std::cos(wrap_pi(t_shift - tc)), std::cos(wrap_pi(-t_shift - tc)));
  }
  if (carbonyl && g.dist(b.h, c) > g.dist(b.h, o[0])) tcos = 0.0;
  if (o[1] == o[2] || b.h == o[2]) tcos = 1.0;
  return {ac, tcos, true};
}

// DH+ energy of one X-H...Y (kcal/mol), PM7 form.
double dh_plus(const Geometry& g, const HBond& b) {
  if (b.skip) return 0.0;
  const double acos_xhy = -std::cos(g.angle(b.x, b.h, b.y));
  if (acos_xhy <= 0.0) return 0.0;
  const EndTerms ex = end_terms(g, b, b.x, b.nx, b.nbx, true);
  if (!ex.ok || ex.torsion_cos < 0.0) return 0.0;
  EndTerms ey = end_terms(g, b, b.y, b.ny, b.nby, false);
  if (!ey.ok) return 0.0;
  ey.torsion_cos = std::fabs(ey.torsion_cos);
  // N and O strengths (a0^2-scaled hartree); PM7 gives sp2 and sp3 centres
  // the same value
  auto strength = [](int z) { return (z == 7 ? -0.171271 : -0.098822) * kA0 * kA0; };
  const double scale = 0.5 * (strength(g.z[b.x]) + strength(g.z[b.y]));
  const double hx = g.dist(b.h, b.x), hy = g.dist(b.h, b.y);
  const double near = std::min(hx, hy);
  double damp = (std::max(hx, hy) - near > 0.5) ? 1.0 - 1.0 / (1.0 + std::exp(-60.0 * (near / 1.2 - 1.0))) : 1.0;
  const double rxy = g.dist(b.x, b.y);
  damp = damp / (1.0 + std::exp(-100.0 * (rxy / 2.4 - 1.0)));
  damp = damp * (1.0 - 1.0 / (1.0 + std::exp(-10.0 * (rxy / 7.0 - 1.0))));
  const double prod = ex.angle_cos * ex.torsion_cos * ey.angle_cos * ey.torsion_cos;
  double e = scale / (rxy * rxy) * acos_xhy * acos_xhy * (1.0 - (1.0 - prod) * (1.0 - prod)) * kEV * kFPC9 * damp;
  if (g.z[b.x] == 8 && g.z[b.y] == 8) {
    // extra stabilization as O-O approaches the short strong-bond limit
    const double dr = std::max(rxy - 2.67, 0.0);
    e += -2.5 * std::exp(-80.0 * dr * dr) * std::pow(acos_xhy, 4);
  }
  return e;
}

}  // namespace

extern "C" {

int mhip_pm7_corrections(int numat, const int* nat, const double* coord, double* e_disp, double* e_hbond,
                         int* n_hbonds) {
  if (numat < 0 || (numat > 0 && (!nat || !coord)) || !e_disp || !e_hbond) return -1;
  for (int i = 0; i < numat; ++i)
    if (!supported(nat[i])) return -1;
  const Geometry g{numat, nat, coord};
  const std::vector<int> nb = bond_counts(g);
  *e_disp = dispersion(g, nb);
  double eh = 0.0;
  int strong = 0;
  std::vector<HBond> hb = find_hbonds(g);
  for (HBond& b : hb) {
    orient(g, b);
    const double e = dh_plus(g, b);
    eh += e;
    if (e < -1.0) ++strong;
  }
  *e_hbond = eh;
  if (n_hbonds) *n_hbonds = strong;
  return 0;
}

}  // extern "C"
