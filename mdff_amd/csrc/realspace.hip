// realspace.hip -- short-range + real-space Ewald pair kernels for gfx950.
//
// One group of LPA lanes per atom walks that atom's full neighbour row
// (every partner j != i within max cut + skin), so each unordered pair is
// evaluated twice, once from each side.  Every quantity is accumulated only
// into atom i (no atomics, deterministic order); pair energies, virials and
// stresses are halved at the end.  The group reduces its partial sums with
// cross-lane shuffles and writes one row of per-atom outputs.
//
// Reference routines:
//   engforce_bmhftd_pbc  src/non_bonded.f90:1076-1228 (BMHFT, BMHFTD, Buckingham)
//   engforce_nmlj_pbc    src/non_bonded.f90:660-827
//   engforce_morse_pbc   src/non_bonded.f90:1241-1390 (energy; analytic force)
//   multipole_ES_dir     src/ewald_dir.f90:40-798 (q-q, q-mu with TT damping, mu-mu)
// Precision: the minimum image and the cutoff tests are FP64 in the
// reference's order; the pair functions run in T (double for the FP64 form,
// float for the mixed form); every per-atom accumulator is FP64.
#include <algorithm>

#include "common.h"

namespace {

constexpr int NPA = 15;  // per-atom columns: u_sr, vir_sr, tau_sr(6), u_cl, tau_cl(6)

struct PairArgs {
  int n, ntype, sr_model, lbmhftd, trunc;
  int qq, qd, dd, do_field, want_stress;
  // exact_image: Fortran nint for the minimum image (ties away from zero);
  // otherwise rint, which equals it for every listed pair when every
  // perpendicular width exceeds twice the list cutoff (no listed pair can sit
  // at a half-integer fractional offset)
  int exact_image;
  // ortho: the cell matrix is diagonal, so r = diag(A) s (the off-diagonal
  // products of the general form are exact zeros: the same bits)
  int ortho;
  double a11, a21, a31, a12, a22, a32, a13, a23, a33;
  double cutsq_c;                    // Coulomb cutoff^2 (d2 > cutsq cycles, src/ewald_dir.f90:202)
  double alpha;
  const AtomRec* rec;                // cell-sorted atoms: fractional coordinates, charge
  const int2* meta;                  // (type, original index) in the same order
  const double* mu;                  // 3n AoS (may be null when no dipoles)
  const PairParam* pp;
  const int32_t* row;                // row starts (n)
  const int32_t* cnt;                // row lengths (n)
  // seg 0: whole rows; 1 (mixed form): the leading cnt_in[i] entries of a
  // row (its inner entries) in FP64, the rest in FP32
  const int32_t* cnt_in;
  int seg;
  double coul;                       // coul_unit
  const int32_t* nl_j;               // sorted positions of the neighbours
  double* fsr;                       // 3n AoS out: f_sr + coul_unit f_dir
  double* fcl;                       // unused by the force kernel
  double* ef;                        // 3n AoS out (field at i)
  double* peratom;                   // NPA x n (column-major: col * n + i)
};

template <typename T>
struct V3 {
  T x, y, z;
};

// d and 1/d from r^2: d correctly rounded (the short-range exponentials are
// sensitive to it); 1/d in FP32 from the hardware reciprocal (within 1 ulp),
// in FP64 by a division
template <typename T>
__device__ __forceinline__ void dist_inv(T r2, T& d, T& dm1) {
  if constexpr (sizeof(T) == 4) {
    d = sqrtf(r2);
    dm1 = __builtin_amdgcn_rcpf(d);
  } else {
    d = sqrt(r2);
    dm1 = T(1) / d;
  }
}

template <int LPA>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = LPA / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPA);
  return v;
}

// The lane's sums over its share (entries r0 + sub, r0 + sub + LPA, ...) of
// one row segment [r0, r1): pair functions in T, forces and fields summed in
// FP64 per pair, energies and stresses in T over the lane's share
template <typename T>
struct RowSums {
  double f0 = 0, f1 = 0, f2 = 0, e0 = 0, e1 = 0, e2 = 0;
  T usr = 0, ucl = 0;
  T ts[6] = {0, 0, 0, 0, 0, 0}, tc[6] = {0, 0, 0, 0, 0, 0};
};

// atom i of a row: fractional coordinates, charge, type, dipole
struct AtomI {
  double si, ti, ui, q, m0, m1, m2;
  int ita;
  bool dip;
};

template <typename T, bool DIP, int LPA, int SRM, bool PREF = true>
__device__ __forceinline__ void pair_row(const PairArgs& a, const PairParam* spp, const AtomI& ai, int r0, int r1,
                                         int sub, RowSums<T>& s) {
  const double si = ai.si, ti = ai.ti, ui = ai.ui;
  const int ita = ai.ita;
  const T qi = (T)ai.q;
  const T alpha = (T)a.alpha, a2 = alpha * alpha;
  const T isqpi = T(0.56418958354775628695);  // 1/sqrt(pi)
  const V3<T> mui{(T)ai.m0, (T)ai.m1, (T)ai.m2};
  const bool dip_i = ai.dip;
  const T coul = (T)a.coul;
  const int srm = SRM;   // short-range model, a compile-time constant of the kernel
  // PREF: two-stage prefetch of the neighbour chain: the row entry two steps
  // ahead and the record of the next entry are in flight while a pair is
  // computed (short segments: each entry loaded when it is reached)
  int pj_next = (PREF && r0 + sub + LPA < r1) ? a.nl_j[r0 + sub + LPA] : 0;
  AtomRec rj_next{};
  int2 mj_next{};
  if (PREF && r0 + sub < r1) {
    const int pj0 = a.nl_j[r0 + sub];
    rj_next = a.rec[pj0];
    mj_next = a.meta[pj0];
  }
  for (int e = r0 + sub; e < r1; e += LPA) {
    AtomRec rj;
    int2 mj;
    if (PREF) {
      rj = rj_next;
      mj = mj_next;
      if (e + LPA < r1) {
        rj_next = a.rec[pj_next];
        mj_next = a.meta[pj_next];
        if (e + 2 * LPA < r1) pj_next = a.nl_j[e + 2 * LPA];
      }
    } else {
      const int pj = a.nl_j[e];
      rj = a.rec[pj];
      mj = a.meta[pj];
    }
    const int j = mj.y;
    double dxs = si - rj.x, dys = ti - rj.y, dzs = ui - rj.z;
    if (a.exact_image) {
      dxs -= nint_d(dxs);
      dys -= nint_d(dys);
      dzs -= nint_d(dzs);
    } else {
      dxs -= rint(dxs);
      dys -= rint(dys);
      dzs -= rint(dzs);
    }
    double rxd, ryd, rzd;
    if (a.ortho) {
      rxd = dxs * a.a11;
      ryd = dys * a.a22;
      rzd = dzs * a.a33;
    } else {
      rxd = dxs * a.a11 + dys * a.a12 + dzs * a.a13;
      ryd = dxs * a.a21 + dys * a.a22 + dzs * a.a23;
      rzd = dxs * a.a31 + dys * a.a32 + dzs * a.a33;
    }
    const double r2d = rxd * rxd + ryd * ryd + rzd * rzd;
    const int jta = mj.x;
    const PairParam& P = spp[ita + a.ntype * jta];
    const bool in_sr = srm != 0 && r2d < P.rcutsq && !(srm == 2 && P.A == 0.0);
    const bool in_cl = (a.qq | a.qd | a.dd) && !(r2d > a.cutsq_c);
    if (!in_sr && !in_cl) continue;
    const T rx = (T)rxd, ry = (T)ryd, rz = (T)rzd, r2 = (T)r2d;
    T d, dm1;
    dist_inv(r2, d, dm1);
    const T dm2 = dm1 * dm1;
    T pfx = T(0), pfy = T(0), pfz = T(0);
    if (in_sr) {
      T w = T(0), u = T(0);
      if (srm == 2) {
        const T erh = (T)P.A * exp(-(T)P.B * d);
        T ir6 = dm2 * dm2 * dm2;
        T ir8 = ir6 * dm2;
        ir6 *= (T)P.C;
        ir8 *= (T)P.D;
        T ir6d = ir6, ir8d = ir8, fd6 = T(0), fd8 = T(0);
        if (a.lbmhftd) {
          T f6, f8;
          tt_damp<T>((T)P.BD, T(1), d, 6, f6, fd6);
          tt_damp<T>((T)P.BD, T(1), d, 8, f8, fd8);
          ir6d = ir6 * f6;
          ir8d = ir8 * f8;
        }
        u = erh - ir6d - ir8d;
        w = ((T)P.B * erh - T(6) * ir6d * dm1 - T(8) * ir8d * dm1 + ir6 * fd6 + ir8 * fd8) * dm1;
      } else if (srm == 1) {
        const T sr2 = (T)P.sigsq * dm2;
        const T srp = pow(sr2, (T)(0.5 * P.pl));
        const T srq = pow(sr2, (T)(0.5 * P.ql));
        u = (T)P.epsp * ((T)P.pl * srq - (T)P.ql * srp);
        if (a.trunc == 1) u -= (T)P.uc;
        if (a.trunc == 2) u += (T)P.uc1 * r2 - (T)P.uc2;
        w = (T)P.fc * (srq - srp) * sr2;
      } else {
        const T erh = exp(-d * (T)P.rhomor);
        const T x = erh * (T)P.rs - T(1);
        u = (T)P.epsmor * (x * x - T(1));
        w = T(2) * (T)P.epsmor * (T)P.rhomor * erh * (T)P.rs * x * dm1;
      }
      const T fx = w * rx, fy = w * ry, fz = w * rz;
      pfx = fx;
      pfy = fy;
      pfz = fz;
      s.usr += u;
      s.ts[0] += (rx * fx);
      s.ts[1] += (ry * fy);
      s.ts[2] += (rz * fz);
      s.ts[3] += (rx * fy);
      s.ts[4] += (rx * fz);
      s.ts[5] += (ry * fz);
    }
    if (in_cl) {
      const T qj = (T)rj.q;
      const T expon = exp(-a2 * r2) * isqpi;
      const T F0 = errfc_pos<T>(alpha * d);
      const T F1 = F0 + T(2) * alpha * d * expon;
      const T dm3 = dm1 * dm2;
      const T F1_dm3 = F1 * dm3;
      T fij0 = T(0), fij1 = T(0), fij2 = T(0), u = T(0);
      const PairParam& Q = P;
      const bool ldamp = Q.ldamp != 0;
      T fdamp = T(1), fdd = T(0), fdamp2 = T(1), fdd2 = T(0);
      if (ldamp) {
        tt_damp<T>((T)Q.db_i, (T)Q.dc_i, d, Q.dk_i, fdamp, fdd);
        if (DIP) tt_damp<T>((T)Q.db_j, (T)Q.dc_j, d, Q.dk_j, fdamp2, fdd2);
      }
      if (a.qq) {
        const T qij = qi * qj;
        u += qij * dm1 * F0;
        // fij = qij T1, T1 = -r dm3 F1 (src/ewald_dir.f90:252-257,528-553)
        fij0 -= qij * F1_dm3 * rx;
        fij1 -= qij * F1_dm3 * ry;
        fij2 -= qij * F1_dm3 * rz;
        if (a.do_field) {
          T g = qj * F1_dm3;
          if (ldamp) g -= qj * (T(1) - fdamp) * dm3;
          s.e0 += (double)(g * rx);
          s.e1 += (double)(g * ry);
          s.e2 += (double)(g * rz);
        }
      }
      if (DIP) {
        V3<T> muj{T(0), T(0), T(0)};
        double m0 = a.mu[3 * j], m1 = a.mu[3 * j + 1], m2 = a.mu[3 * j + 2];
        const bool dip_j = (m0 != 0.0 || m1 != 0.0 || m2 != 0.0);
        muj = {(T)m0, (T)m1, (T)m2};
        if ((a.qd && (dip_i || dip_j)) || (a.dd && dip_i && dip_j)) {
          const T d3 = r2 * d;
          const T F2 = F1 + T(4.0 / 3.0) * alpha * a2 * d3 * expon;
          const T dm5 = dm3 * dm2;
          const T F2_dm5 = T(3) * F2 * dm5;
          const T rmi = rx * mui.x + ry * mui.y + rz * mui.z;
          const T rmj = rx * muj.x + ry * muj.y + rz * muj.z;
          if (a.qd && (dip_i || dip_j)) {
            // v = qj mu_i - qi mu_j (src/ewald_dir.f90:599-644)
            const T vx = qj * mui.x - qi * muj.x, vy = qj * mui.y - qi * muj.y, vz = qj * mui.z - qi * muj.z;
            const T rv = rx * vx + ry * vy + rz * vz;
            u -= dm3 * F1 * rv;
            fij0 += F2_dm5 * rx * rv - F1_dm3 * vx;
            fij1 += F2_dm5 * ry * rv - F1_dm3 * vy;
            fij2 += F2_dm5 * rz * rv - F1_dm3 * vz;
            if (ldamp) {
              const T F1d = T(1) - fdamp, F2d = F1d + (d / T(3)) * fdd;
              const T F1d2 = T(1) - fdamp2, F2d2 = F1d2 + (d / T(3)) * fdd2;
              u -= dm3 * (qi * F1d2 * rmj - qj * F1d * rmi);
              const T ci = T(3) * F2d * dm5 * rmi, cj = T(3) * F2d2 * dm5 * rmj;
              const T di = F1d * dm3, dj = F1d2 * dm3;
              fij0 += -qj * (ci * rx - di * mui.x) + qi * (cj * rx - dj * muj.x);
              fij1 += -qj * (ci * ry - di * mui.y) + qi * (cj * ry - dj * muj.y);
              fij2 += -qj * (ci * rz - di * mui.z) + qi * (cj * rz - dj * muj.z);
            }
          }
          if (a.dd && dip_i && dip_j) {
            // mu-mu (src/ewald_dir.f90:557-593), T3 contraction written out
            const T mm = mui.x * muj.x + mui.y * muj.y + mui.z * muj.z;
            u -= F2_dm5 * rmi * rmj - F1_dm3 * mm;
            if (a.do_field) {
              s.e0 += (double)(F2_dm5 * rx * rmj - F1_dm3 * muj.x);
              s.e1 += (double)(F2_dm5 * ry * rmj - F1_dm3 * muj.y);
              s.e2 += (double)(F2_dm5 * rz * rmj - F1_dm3 * muj.z);
            }
            const T d5 = d3 * r2;
            const T F3 = F2 + T(8.0 / 15.0) * alpha * a2 * a2 * d5 * expon;
            const T F3_dm7 = T(15) * F3 * dm5 * dm2;
            const T s3 = -F3_dm7 * rmi * rmj;
            fij0 -= s3 * rx + F2_dm5 * (mui.x * rmj + mm * rx + rmi * muj.x);
            fij1 -= s3 * ry + F2_dm5 * (mui.y * rmj + mm * ry + rmi * muj.y);
            fij2 -= s3 * rz + F2_dm5 * (mui.z * rmj + mm * rz + rmi * muj.z);
          }
        }
      }
      pfx -= coul * fij0;
      pfy -= coul * fij1;
      pfz -= coul * fij2;
      s.ucl += u;
      s.tc[0] -= (T(2) * rx * fij0);
      s.tc[1] -= (T(2) * ry * fij1);
      s.tc[2] -= (T(2) * rz * fij2);
      s.tc[3] -= (rx * fij1 + ry * fij0);
      s.tc[4] -= (rx * fij2 + rz * fij0);
      s.tc[5] -= (ry * fij2 + rz * fij1);
    }
    s.f0 += (double)pfx;
    s.f1 += (double)pfy;
    s.f2 += (double)pfz;
  }
}

// seg 0: the whole row in T; seg 1 (T = float): the row's leading
// cnt_in[i] entries (its inner entries) in FP64, the rest in FP32, one pass.
// The lane's FP64 and FP32 sums are added in FP64 before the group sums.
template <typename T, bool DIP, int LPA, int SRM>
__global__ void __launch_bounds__(256) k_pair_force(PairArgs a) {
  __shared__ PairParam spp[64];
  for (int k = threadIdx.x; k < a.ntype * a.ntype; k += blockDim.x) spp[k] = a.pp[k];
  __syncthreads();
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int pi = gtid / LPA;             // atoms in cell order: neighbouring groups share neighbours
  const int sub = gtid % LPA;
  if (pi >= a.n) return;  // whole groups exit together
  const AtomRec ri = a.rec[pi];
  const int2 mi = a.meta[pi];
  const int i = mi.y;
  AtomI ai{ri.x, ri.y, ri.z, ri.q, 0.0, 0.0, 0.0, mi.x, false};
  if (DIP) {
    ai.m0 = a.mu[3 * i];
    ai.m1 = a.mu[3 * i + 1];
    ai.m2 = a.mu[3 * i + 2];
    ai.dip = (ai.m0 != 0.0 || ai.m1 != 0.0 || ai.m2 != 0.0);
  }
  int r0 = a.row[i];
  const int r1 = r0 + a.cnt[i];
  // energies and stresses: summed in T over the lane's share of the row,
  // then in FP64 across the lanes and atoms (each lane holds a few dozen
  // pairs, so the T rounding of these global sums stays far below the
  // force error of the mixed form)
  double f0 = 0, f1 = 0, f2 = 0, e0 = 0, e1 = 0, e2 = 0;
  double usr = 0, ucl = 0, ts[6] = {0, 0, 0, 0, 0, 0}, tc[6] = {0, 0, 0, 0, 0, 0};
  if constexpr (sizeof(T) == 4) {
    if (a.seg == 1) {
      const int rm = r0 + a.cnt_in[i];
      RowSums<double> sd;
      pair_row<double, DIP, LPA, SRM, false>(a, spp, ai, r0, rm, sub, sd);
      f0 = sd.f0; f1 = sd.f1; f2 = sd.f2;
      e0 = sd.e0; e1 = sd.e1; e2 = sd.e2;
      usr = sd.usr; ucl = sd.ucl;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        ts[k] = sd.ts[k];
        tc[k] = sd.tc[k];
      }
      r0 = rm;
    }
  }
  {
    RowSums<T> st;
    pair_row<T, DIP, LPA, SRM>(a, spp, ai, r0, r1, sub, st);
    f0 += st.f0; f1 += st.f1; f2 += st.f2;
    e0 += st.e0; e1 += st.e1; e2 += st.e2;
    usr += (double)st.usr;
    ucl += (double)st.ucl;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      ts[k] += (double)st.ts[k];
      tc[k] += (double)st.tc[k];
    }
  }
  f0 = group_sum<LPA>(f0);
  f1 = group_sum<LPA>(f1);
  f2 = group_sum<LPA>(f2);
  if (a.do_field) {
    e0 = group_sum<LPA>(e0);
    e1 = group_sum<LPA>(e1);
    e2 = group_sum<LPA>(e2);
  }
  double cols[NPA];
  if (a.want_stress) {
    cols[0] = group_sum<LPA>(usr);
    cols[1] = 0.0;                        // set from the stress trace below
#pragma unroll
    for (int k = 0; k < 6; ++k) cols[2 + k] = group_sum<LPA>(ts[k]);
    cols[1] = (cols[2] + cols[3]) + cols[4];  // virial w r^2 = trace of r (w r)
    cols[8] = group_sum<LPA>(ucl);
#pragma unroll
    for (int k = 0; k < 6; ++k) cols[9 + k] = group_sum<LPA>(tc[k]);
  }
  if (sub == 0) {
    a.fsr[3 * i] = f0;
    a.fsr[3 * i + 1] = f1;
    a.fsr[3 * i + 2] = f2;
    if (a.do_field) {
      a.ef[3 * i] = e0;
      a.ef[3 * i + 1] = e1;
      a.ef[3 * i + 2] = e2;
    }
    if (a.want_stress) {
#pragma unroll
      for (int k = 0; k < NPA; ++k) a.peratom[(size_t)k * a.n + i] = cols[k];
    }
  }
}

// Field-only pass of the SCF (src/moment_from_pola.f90:327-330, 401-404).
// mode 0: static field of the charges, TT-damped; mode 1: field of the
// dipoles (mu-mu, undamped, only pairs where both dipoles are non-zero).
template <typename T, int MODE, int LPA>
__global__ void __launch_bounds__(256) k_pair_field(PairArgs a) {
  __shared__ PairParam spp[64];
  for (int k = threadIdx.x; k < a.ntype * a.ntype; k += blockDim.x) spp[k] = a.pp[k];
  __syncthreads();
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int pi = gtid / LPA;
  const int sub = gtid % LPA;
  if (pi >= a.n) return;
  const AtomRec ri = a.rec[pi];
  const int2 mi = a.meta[pi];
  const int i = mi.y;
  const double si = ri.x, ti = ri.y, ui = ri.z;
  const int ita = mi.x;
  const T alpha = (T)a.alpha, a2 = alpha * alpha;
  const T isqpi = T(0.56418958354775628695);
  bool dip_i = true;
  if (MODE == 1) dip_i = (a.mu[3 * i] != 0.0 || a.mu[3 * i + 1] != 0.0 || a.mu[3 * i + 2] != 0.0);
  double e0 = 0, e1 = 0, e2 = 0;
  const int r0 = a.row[i], r1 = r0 + a.cnt[i];
  if (dip_i) {
    // neighbour chain prefetched two stages ahead, as in k_pair_force
    int pj_next = (r0 + sub + LPA < r1) ? a.nl_j[r0 + sub + LPA] : 0;
    AtomRec rj_next{};
    int2 mj_next{};
    if (r0 + sub < r1) {
      const int pj0 = a.nl_j[r0 + sub];
      rj_next = a.rec[pj0];
      mj_next = a.meta[pj0];
    }
    for (int e = r0 + sub; e < r1; e += LPA) {
      const AtomRec rj = rj_next;
      const int2 mj = mj_next;
      if (e + LPA < r1) {
        rj_next = a.rec[pj_next];
        mj_next = a.meta[pj_next];
        if (e + 2 * LPA < r1) pj_next = a.nl_j[e + 2 * LPA];
      }
      const int j = mj.y;
      T m0, m1, m2;
      if (MODE == 1) {
        double d0 = a.mu[3 * j], d1 = a.mu[3 * j + 1], d2 = a.mu[3 * j + 2];
        if (d0 == 0.0 && d1 == 0.0 && d2 == 0.0) continue;
        m0 = (T)d0; m1 = (T)d1; m2 = (T)d2;
      }
      double dxs = si - rj.x, dys = ti - rj.y, dzs = ui - rj.z;
      if (a.exact_image) {
        dxs -= nint_d(dxs);
        dys -= nint_d(dys);
        dzs -= nint_d(dzs);
      } else {
        dxs -= rint(dxs);
        dys -= rint(dys);
        dzs -= rint(dzs);
      }
      double rxd, ryd, rzd;
      if (a.ortho) {
        rxd = dxs * a.a11;
        ryd = dys * a.a22;
        rzd = dzs * a.a33;
      } else {
        rxd = dxs * a.a11 + dys * a.a12 + dzs * a.a13;
        ryd = dxs * a.a21 + dys * a.a22 + dzs * a.a23;
        rzd = dxs * a.a31 + dys * a.a32 + dzs * a.a33;
      }
      const double r2d = rxd * rxd + ryd * ryd + rzd * rzd;
      if (r2d > a.cutsq_c) continue;
      const T rx = (T)rxd, ry = (T)ryd, rz = (T)rzd, r2 = (T)r2d;
      T d, dm1;
      dist_inv(r2, d, dm1);
      const T dm2 = dm1 * dm1, dm3 = dm1 * dm2;
      const T expon = exp(-a2 * r2) * isqpi;
      const T F0 = errfc_pos<T>(alpha * d);
      const T F1 = F0 + T(2) * alpha * d * expon;
      if (MODE == 0) {
        const T qj = (T)rj.q;
        T g = qj * F1 * dm3;
        const PairParam& P = spp[ita + a.ntype * mj.x];
        if (P.ldamp) {
          T fdamp, fdd;
          tt_damp<T>((T)P.db_i, (T)P.dc_i, d, P.dk_i, fdamp, fdd);
          g -= qj * (T(1) - fdamp) * dm3;
        }
        e0 += (double)(g * rx);
        e1 += (double)(g * ry);
        e2 += (double)(g * rz);
      } else {
        const T F2 = F1 + T(4.0 / 3.0) * alpha * a2 * r2 * d * expon;
        const T F2_dm5 = T(3) * F2 * dm3 * dm2;
        const T F1_dm3 = F1 * dm3;
        const T rm = rx * m0 + ry * m1 + rz * m2;
        e0 += (double)(F2_dm5 * rx * rm - F1_dm3 * m0);
        e1 += (double)(F2_dm5 * ry * rm - F1_dm3 * m1);
        e2 += (double)(F2_dm5 * rz * rm - F1_dm3 * m2);
      }
    }
  }
  e0 = group_sum<LPA>(e0);
  e1 = group_sum<LPA>(e1);
  e2 = group_sum<LPA>(e2);
  if (sub == 0) {
    a.ef[3 * i] = e0;
    a.ef[3 * i + 1] = e1;
    a.ef[3 * i + 2] = e2;
  }
}

// Real-space part of the electric field gradient at each atom
// (src/ewald_dir.f90:543-546, 577-582), FP64 in both precision forms: the
// q-q task adds -q_j T2(r_ij), the mu-mu task T3(r_ij).mu_j for pairs where
// both dipoles are non-zero (the reference's dd block, :557; its q-mu block
// adds nothing to the gradient).  Every pair is seen from both atoms; the
// reference's j-side term -T3(r_ij).mu_i is T3(r_ji).mu_i (T3 is odd).
// out: 6 per atom, (xx, yy, zz, xy, xz, yz).
template <int LPA>
__global__ void __launch_bounds__(256) k_pair_efg(PairArgs a, double* __restrict__ out) {
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int pi = gtid / LPA;
  const int sub = gtid % LPA;
  if (pi >= a.n) return;
  const AtomRec ri = a.rec[pi];
  const int i = a.meta[pi].y;
  const double alpha = a.alpha, a2 = alpha * alpha, a3 = a2 * alpha, a5 = a3 * a2;
  const double isqpi = 0.56418958354775628695;
  const bool dd = a.dd && a.mu != nullptr &&
                  (a.mu[3 * i] != 0.0 || a.mu[3 * i + 1] != 0.0 || a.mu[3 * i + 2] != 0.0);
  double g[6] = {0, 0, 0, 0, 0, 0};
  const int r0 = a.row[i], r1 = r0 + a.cnt[i];
  for (int e = r0 + sub; e < r1; e += LPA) {
    const int pj = a.nl_j[e];
    const AtomRec rj = a.rec[pj];
    const int j = a.meta[pj].y;
    double dxs = ri.x - rj.x, dys = ri.y - rj.y, dzs = ri.z - rj.z;
    if (a.exact_image) {
      dxs -= nint_d(dxs);
      dys -= nint_d(dys);
      dzs -= nint_d(dzs);
    } else {
      dxs -= rint(dxs);
      dys -= rint(dys);
      dzs -= rint(dzs);
    }
    const double rx = dxs * a.a11 + dys * a.a12 + dzs * a.a13;
    const double ry = dxs * a.a21 + dys * a.a22 + dzs * a.a23;
    const double rz = dxs * a.a31 + dys * a.a32 + dzs * a.a33;
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 > a.cutsq_c) continue;
    const double d = sqrt(r2), dm1 = 1.0 / d, dm3 = dm1 / r2, dm5 = dm3 / r2, dm7 = dm5 / r2;
    const double d3 = r2 * d, d5 = d3 * r2;
    const double expon = exp(-a2 * r2) * isqpi;
    const double F0 = errfc_pos<double>(alpha * d);
    const double F1 = F0 + 2.0 * alpha * d * expon;
    const double F2 = F1 + 4.0 * a3 * d3 * expon / 3.0;
    const double F2_dm5 = 3.0 * F2 * dm5;
    const double F1_dm3 = F1 * dm3;
    const double r[3] = {rx, ry, rz};
    constexpr int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    if (a.qq) {
      const double qj = rj.q;
#pragma unroll
      for (int x = 0; x < 6; ++x) {
        double t2 = r[ia[x]] * r[ib[x]] * F2_dm5;
        if (x < 3) t2 -= F1_dm3;
        g[x] -= qj * t2;
      }
    }
    if (dd) {
      const double m[3] = {a.mu[3 * j], a.mu[3 * j + 1], a.mu[3 * j + 2]};
      if (m[0] != 0.0 || m[1] != 0.0 || m[2] != 0.0) {
        const double F3 = F2 + 8.0 * a5 * d5 * expon / 15.0;
        const double F3_dm7 = 15.0 * F3 * dm7;
        const double rm = r[0] * m[0] + r[1] * m[1] + r[2] * m[2];
#pragma unroll
        for (int x = 0; x < 6; ++x) {
          const int p = ia[x], q = ib[x];
          double t3 = -r[p] * r[q] * rm * F3_dm7 + (m[p] * r[q] + r[p] * m[q]) * F2_dm5;
          if (x < 3) t3 += rm * F2_dm5;
          g[x] += t3;
        }
      }
    }
  }
#pragma unroll
  for (int x = 0; x < 6; ++x) g[x] = group_sum<LPA>(g[x]);
  if (sub == 0) {
#pragma unroll
    for (int x = 0; x < 6; ++x) out[6 * (size_t)i + x] = g[x];
  }
}

PairArgs make_args(mdff_ctx* c) {
  PairArgs a{};
  a.n = c->n;
  a.ntype = c->ntype;
  a.sr_model = c->sr_model;
  a.lbmhftd = c->d.lbmhftd;
  a.trunc = c->d.trunc;
  const double* A = c->A;
  a.a11 = A[0]; a.a21 = A[1]; a.a31 = A[2];
  a.a12 = A[3]; a.a22 = A[4]; a.a32 = A[5];
  a.a13 = A[6]; a.a23 = A[7]; a.a33 = A[8];
  a.ortho = (A[1] == 0.0 && A[2] == 0.0 && A[3] == 0.0 && A[5] == 0.0 && A[6] == 0.0 && A[7] == 0.0) ? 1 : 0;
  a.cutsq_c = c->rc_coul * c->rc_coul;
  {
    const double rlist = std::max(c->rc_vdw, c->rc_coul) + c->d.skindiff;
    a.exact_image = 0;
    for (int d = 0; d < 3; ++d)
      if (!(c->W[d] > 2.0 * rlist * (1.0 + 1e-6))) a.exact_image = 1;
  }
  a.alpha = c->alpha;
  a.rec = c->arec;
  a.meta = c->ameta;
  a.pp = c->d_pp;
  a.row = c->nl_start;
  a.cnt = c->nl_count;
  a.coul = c->d.coul_unit;
  a.nl_j = c->nl_j;
  return a;
}

template <typename T, bool DIP>
void launch_force(mdff_ctx* c, const PairArgs& a) {
  constexpr int LPA = 16;
  dim3 grid(div_up((int64_t)c->n * LPA, 256));
  // one kernel per short-range model (0 none, 1 n-m LJ, 2 BMH family, 3 Morse)
  switch (a.sr_model) {
    case 1:
      PROF_LAUNCH(c, PK_PAIR, hipLaunchKernelGGL((k_pair_force<T, DIP, LPA, 1>), grid, dim3(256), 0, c->stream, a));
      break;
    case 2:
      PROF_LAUNCH(c, PK_PAIR, hipLaunchKernelGGL((k_pair_force<T, DIP, LPA, 2>), grid, dim3(256), 0, c->stream, a));
      break;
    case 3:
      PROF_LAUNCH(c, PK_PAIR, hipLaunchKernelGGL((k_pair_force<T, DIP, LPA, 3>), grid, dim3(256), 0, c->stream, a));
      break;
    default:
      PROF_LAUNCH(c, PK_PAIR, hipLaunchKernelGGL((k_pair_force<T, DIP, LPA, 0>), grid, dim3(256), 0, c->stream, a));
      break;
  }
}

template <typename T, int MODE>
void launch_field(mdff_ctx* c, const PairArgs& a) {
  constexpr int LPA = 16;
  dim3 grid(div_up((int64_t)c->n * LPA, 256));
  PROF_LAUNCH(c, PK_FIELD, hipLaunchKernelGGL((k_pair_field<T, MODE, LPA>), grid, dim3(256), 0, c->stream, a));
}

}  // namespace

// Force pass: short range + real-space Ewald with the engforce task set.
// fsr, fcl, ef: 3n AoS outputs; per-atom energy/virial/stress columns go to
// c->peratom.
int realspace_force_impl(mdff_ctx* c, double* fsr, double* fcl, double* ef, bool with_field) {
  PairArgs a = make_args(c);
  a.qq = c->d.lcoulomb ? c->task_qq : 0;
  a.qd = c->d.lcoulomb ? c->task_qd : 0;
  a.dd = c->d.lcoulomb ? c->task_dd : 0;
  a.do_field = with_field ? 1 : 0;
  a.want_stress = 1;
  a.mu = c->mu;
  a.fsr = fsr;
  a.fcl = fcl;
  a.ef = ef;
  a.peratom = c->peratom;
  const bool dip = (a.qd || a.dd);
  const bool fp32 = c->d.precision == MDFF_PREC_MIXED;
  if (fp32) {
    if (c->rc_inner > 0.0) {
      // the inner entries of the rows in FP64, the rest in FP32
      a.cnt_in = c->nl_count_in;
      a.seg = 1;
    }
    if (dip) launch_force<float, true>(c, a);
    else launch_force<float, false>(c, a);
  } else {
    if (dip) launch_force<double, true>(c, a);
    else launch_force<double, false>(c, a);
  }
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

int realspace_field(mdff_ctx* c, int mode, const double* mu_in, double* ef_out, double alpha) {
  PairArgs a = make_args(c);
  if (alpha > 0.0) a.alpha = alpha;   // alphaES of this call only (scf_kO_v2 inner loop)
  a.mu = mu_in;
  a.ef = ef_out;
  a.do_field = 1;
  const bool fp32 = c->d.precision == MDFF_PREC_MIXED;
  if (mode == 0) {
    if (fp32) launch_field<float, 0>(c, a);
    else launch_field<double, 0>(c, a);
  } else {
    if (fp32) launch_field<float, 1>(c, a);
    else launch_field<double, 1>(c, a);
  }
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

// The LJ pass of an LJ + BMHFT(D) evaluation: the n-m LJ pair kernel alone
// (no Coulomb terms), its per-atom energy / virial / stress columns into
// peratom and its forces into f_scratch, which the caller drops.
int realspace_lj_extra(mdff_ctx* c, double* f_scratch, double* peratom) {
  PairArgs a = make_args(c);
  a.sr_model = 1;
  a.want_stress = 1;
  a.fsr = f_scratch;
  a.peratom = peratom;
  if (c->d.precision == MDFF_PREC_MIXED) {
    if (c->rc_inner > 0.0) {
      a.cnt_in = c->nl_count_in;
      a.seg = 1;
    }
    launch_force<float, false>(c, a);
  } else {
    launch_force<double, false>(c, a);
  }
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}

// Real-space electric field gradient of the engforce task set into out
// (6 per atom, FP64); the atoms and lists of the current evaluation.
int realspace_efg(mdff_ctx* c, double* out) {
  PairArgs a = make_args(c);
  a.qq = c->task_qq;
  a.dd = c->task_dd;
  a.mu = c->mu;
  constexpr int LPA = 16;
  hipLaunchKernelGGL((k_pair_efg<LPA>), dim3(div_up((int64_t)c->n * LPA, 256)), dim3(256), 0, c->stream, a, out);
  HIP_TRY(c, hipGetLastError());
  return MDFF_OK;
}
