// BDF stability-limit detection (STALD): the decision function of
// cvSLdet, ThirdParty/sundials/src/cvodes/cvodes.c:8256-8519.
//
// Input: the squared scaled derivative norms of the last five steps taken at
// order q >= 3, ss[i][k] with i = 0 the newest step and k = 0, 1, 2 the
// derivatives of order q - 1, q, q + 1 (cvBDFStab, :8158-8219).  The routine
// estimates the dominant characteristic root rr of the recent step sequence,
// from the ratios of successive norms when the three orders agree (the
// normal-matrix case) or else from the common root of three quartics, refined
// by up to three Newton corrections; it then checks rr against the variance
// relations of the BDF method of order q.  Returns 1, 2 or 3 for a stable
// root, 4, 5 or 6 when rr exceeds 0.98 (a stability-limit violation: the
// caller lowers the order), and a negative value when no decision is
// possible (-1 data range too wide, -2 / -3 inconsistent normal case, -4 /
// -5 quartic elimination failed, -6 no Newton correction converged, -7 / -8
// / -9 the variance check failed).
//
// Plain scalar code, the same on every lane of a team (uniform inputs); also
// compiled for the host, where the tests compare it with CVODES's routine.
// The arithmetic is arranged for a short dependent chain on the GPU: ratios
// that only meet a threshold or each other are formed with reciprocals
// computed once, the two divisions of each variance correction are merged,
// and the remaining quotients use a refined hardware reciprocal on the
// device (a few units in the last place; every decision of the detector is a
// comparison with a threshold that rounding at that level does not move in
// practice, which the tests check against CVODES's exact arithmetic).
#pragma once

#if defined(__HIPCC__)
#define AMX_HD __host__ __device__
#else
#define AMX_HD
#endif

namespace amx {

// a / b: two Newton steps on the hardware reciprocal on the device, the
// division itself on the host
AMX_HD inline double stald_div(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
  r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
  return a * r;
#else
  return a / b;
#endif
}

AMX_HD inline int stald_flag(const double (&ss)[5][3], int q) {
  constexpr double TINY_ = 1.0e-10;
  constexpr double RRCUT = 0.98, VRRTOL = 1.0e-4, VRRT2 = 5.0e-4, SQTOL = 1.0e-3,
                   RRTOL = 1.0e-2;
  double smax[3], ssmax[3], rssmax[3], rav[3], vrat[3];
  double qc[5][3];  // quartic coefficients: qc[d][k] multiplies rr^d (d = 0..4; no rr^2 term)
  // minimum / maximum with the operand order of SUNMIN / SUNMAX
  auto mn = [](double a, double b) { return a < b ? a : b; };
  auto mx = [](double a, double b) { return a > b ? a : b; };
  for (int k = 0; k < 3; ++k) {
    double smin = ss[0][k], smx = 0.0;
    for (int i = 0; i < 5; ++i) {
      smin = mn(smin, ss[i][k]);
      smx = mx(smx, ss[i][k]);
    }
    if (smin < TINY_ * smx) return -1;
    smax[k] = smx;
    ssmax[k] = smx * smx;
    rssmax[k] = stald_div(1.0, ssmax[k]);
    double sr = 0.0, sr2 = 0.0;
    for (int i = 0; i < 4; ++i) {
      const double r = stald_div(ss[i][k], ss[i + 1][k]);
      sr += r;
      sr2 += r * r;
    }
    rav[k] = 0.25 * sr;
    const double v = 0.25 * sr2 - rav[k] * rav[k];
    vrat[k] = v < 0.0 ? -v : v;
    qc[0][k] = ss[0][k] * ss[2][k] - ss[1][k] * ss[1][k];
    qc[1][k] = ss[1][k] * ss[2][k] - ss[0][k] * ss[3][k];
    qc[2][k] = 0.0;
    qc[3][k] = ss[1][k] * ss[4][k] - ss[2][k] * ss[3][k];
    qc[4][k] = ss[3][k] * ss[3][k] - ss[2][k] * ss[4][k];
  }
  auto quartic = [&](int k, double s) {
    return qc[0][k] + s * (qc[1][k] + s * s * (qc[3][k] + s * qc[4][k]));
  };
  auto fabs_ = [](double v) { return v < 0.0 ? -v : v; };

  double rr = 0.0;
  int kflag;
  const double vmin = mn(vrat[0], mn(vrat[1], vrat[2]));
  const double vmax = mx(vrat[0], mx(vrat[1], vrat[2]));
  if (vmin < VRRTOL * VRRTOL) {
    // normal-matrix case: the three ratio sequences share their mean
    if (vmax > VRRT2 * VRRT2) return -2;
    rr = (rav[0] + rav[1] + rav[2]) * (1.0 / 3.0);
    double dmax = 0.0;
    for (int k = 0; k < 3; ++k) dmax = mx(dmax, fabs_(rav[k] - rr));
    if (dmax > VRRT2) return -3;
    kflag = 1;
  } else {
    // eliminate the leading coefficient and then rr^3 between the quartics,
    // leaving a linear equation a4 rr + a0 = 0 in the third
    double e[5][3];
    for (int d = 0; d < 5; ++d)
      for (int k = 0; k < 3; ++k) e[d][k] = qc[d][k];
    if (fabs_(e[4][0]) < TINY_ * ssmax[0]) return -4;
    for (int k = 1; k < 3; ++k) {
      const double f = stald_div(e[4][k], e[4][0]);
      for (int d = 0; d < 4; ++d) e[d][k] -= f * e[d][0];
      e[4][k] = 0.0;
    }
    if (fabs_(e[3][1]) < TINY_ * ssmax[1]) return -4;
    {
      const double f = stald_div(e[3][2], e[3][1]);
      for (int d = 0; d < 3; ++d) e[d][2] -= f * e[d][1];
    }
    if (fabs_(e[1][2]) < TINY_ * ssmax[2]) return -4;
    rr = stald_div(-e[0][2], e[1][2]);
    if (rr < TINY_ || rr > 100.0) return -5;
    double qkr[3];
    double sqmax = 0.0;
    for (int k = 0; k < 3; ++k) {
      qkr[k] = quartic(k, rr);
      const double a = fabs_(qkr[k]) * rssmax[k];
      sqmax = a > sqmax ? a : sqmax;
    }
    if (sqmax < SQTOL) {
      kflag = 2;
    } else {
      // Newton corrections of rr on each quartic; keep the candidate whose
      // largest scaled residual over the three quartics is smallest
      double sqmin = 0.0;
      bool done = false;
      for (int it = 0; it < 3 && !done; ++it) {
        double cand[3];
        for (int k = 0; k < 3; ++k) {
          const double dq = qc[1][k] + rr * rr * (3.0 * qc[3][k] + rr * 4.0 * qc[4][k]);
          const double dr = (fabs_(dq) > TINY_ * ssmax[k]) ? stald_div(-qkr[k], dq) : 0.0;
          cand[k] = rr + dr;
        }
        double res[3][3], worst[3];
        for (int k = 0; k < 3; ++k) {
          double w = 0.0;
          for (int j = 0; j < 3; ++j) {
            res[j][k] = quartic(j, cand[k]);
            const double a = fabs_(res[j][k]) * rssmax[j];
            w = a > w ? a : w;
          }
          worst[k] = w;
        }
        int kmin = 0;
        sqmin = worst[0] + 1.0;
        for (int k = 0; k < 3; ++k)
          if (worst[k] < sqmin) {
            kmin = k;
            sqmin = worst[k];
          }
        rr = cand[kmin];
        if (sqmin < SQTOL) {
          done = true;
        } else {
          for (int j = 0; j < 3; ++j) qkr[j] = res[j][kmin];
        }
      }
      if (sqmin > SQTOL) return -6;
      // (a residual exactly at the tolerance leaves the flag at 0, as in CVODES)
      kflag = done ? 3 : 0;
    }
  }

  // given rr: the variances sigsq[k] of the three orders, and rr recomputed
  // from them, must agree with rr
  double sigsq[3];
  for (int k = 0; k < 3; ++k) {
    const double a = ss[0][k];
    const double b = ss[1][k] * rr;
    const double c = ss[2][k] * rr * rr;
    const double d = ss[3][k] * rr * rr * rr;
    const double d1a = a - b, d1b = b - c, d1c = c - d;
    const double d2a = d1a - d1b, d2b = d1b - d1c;
    const double d3a = d2a - d2b;
    if (fabs_(d1b) < TINY_ * smax[k]) return -7;
    const double cest = stald_div(-d3a, d1b);
    if (cest < TINY_ || cest > 4.0) return -7;
    // (d2b / cest) / rr^2 with cest = -d3a / d1b
    sigsq[k] = ss[2][k] + stald_div(d2b * d1b, -d3a * (rr * rr));
  }
  if (sigsq[1] < TINY_) return -8;
  const double rs1 = stald_div(1.0, sigsq[1]);
  const double ratp = sigsq[2] * rs1;
  const double ratm = sigsq[0] * rs1;
  const double qfac1 = 0.25 * (q * q - 1.0);
  const double qfac2 = q == 3 ? 1.0 : (q == 4 ? 2.0 / 3.0 : 2.0 / (q - 1.0));
  const double bb = ratp * ratm - 1.0 - qfac1 * ratp;
  const double tem = 1.0 - qfac2 * bb;
  if (fabs_(tem) < TINY_) return -8;
  const double rrb = stald_div(1.0, tem);
  if (fabs_(rrb - rr) > RRTOL) return -9;
  if (rr > RRCUT && kflag >= 1) kflag += 3;
  return kflag;
}

}  // namespace amx
