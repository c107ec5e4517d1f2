// Moré-Thuente line search (Moré & Thuente, ACM TOMS 20 (1994) 286-307) with
// the approximate Wolfe acceptance of Hager & Zhang (SIAM J. Optim. 16 (2005)
// 170-192).  Host-side scalar logic shared by the single-lattice and batched
// minimizers.
#pragma once
#include <algorithm>
#include <cmath>

namespace a2c {

// Moré-Thuente line search: the step-selection rule (dcstep) and the driver
// (dcsrch) of the published algorithm, reverse-communication style.
struct MoreThuente {
  double ftol = 1e-4, gtol = 0.9, xtol = 1e-10, stpmin = 1e-20, stpmax = 1e20;
  double epsf = 0.0;  // tolerated rise of f for the approximate Wolfe test
  bool brackt = false;
  int stage = 1;
  double finit = 0, ginit = 0, gtest = 0, width = 0, width1 = 0;
  double stx = 0, fx = 0, gx = 0, sty = 0, fy = 0, gy = 0, stmin = 0, stmax = 0;

  static void cstep(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp,
                    double fp, double dp, bool& brackt, double stpmin, double stpmax) {
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf;
    if (fp > fx) {
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
      double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp < stx) gamma = -gamma;
      const double p = (gamma - dx) + theta, q = ((gamma - dx) + gamma) + dp, r = p / q;
      const double stpc = stx + r * (stp - stx);
      const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
      stpf = (std::fabs(stpc - stx) < std::fabs(stpq - stx)) ? stpc : stpc + (stpq - stpc) / 2.0;
      brackt = true;
    } else if (sgnd < 0.0) {
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
      double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp > stx) gamma = -gamma;
      const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dx, r = p / q;
      const double stpc = stp + r * (stx - stp);
      const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
      stpf = (std::fabs(stpc - stp) > std::fabs(stpq - stp)) ? stpc : stpq;
      brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
      double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
      if (stp > stx) gamma = -gamma;
      const double p = (gamma - dp) + theta, q = (gamma + (dx - dp)) + gamma, r = p / q;
      double stpc;
      if (r < 0.0 && gamma != 0.0) stpc = stp + r * (stx - stp);
      else if (stp > stx) stpc = stpmax;
      else stpc = stpmin;
      const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
      if (brackt) {
        stpf = (std::fabs(stpc - stp) < std::fabs(stpq - stp)) ? stpc : stpq;
        if (stp > stx) stpf = std::min(stp + 0.66 * (sty - stp), stpf);
        else stpf = std::max(stp + 0.66 * (sty - stp), stpf);
      } else {
        stpf = (std::fabs(stpc - stp) > std::fabs(stpq - stp)) ? stpc : stpq;
        stpf = std::min(stpmax, stpf);
        stpf = std::max(stpmin, stpf);
      }
    } else {
      if (brackt) {
        const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
        if (stp > sty) gamma = -gamma;
        const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dy, r = p / q;
        stpf = stp + r * (sty - stp);
      } else if (stp > stx) {
        stpf = stpmax;
      } else {
        stpf = stpmin;
      }
    }
    if (fp > fx) {
      sty = stp; fy = fp; dy = dp;
    } else {
      if (sgnd < 0.0) { sty = stx; fy = fx; dy = dx; }
      stx = stp; fx = fp; dx = dp;
    }
    stp = stpf;
  }

  void start(double f, double g, double stp) {
    brackt = false;
    stage = 1;
    finit = f;
    ginit = g;
    gtest = ftol * ginit;
    width = stpmax - stpmin;
    width1 = 2.0 * width;
    stx = 0.0; fx = finit; gx = ginit;
    sty = 0.0; fy = finit; gy = ginit;
    stmin = 0.0;
    stmax = stp + 4.0 * stp;
  }

  // returns 0: evaluate at the new stp; 1: converged; 2: warning (stop, use stp)
  int iterate(double f, double g, double& stp) {
    const double ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
    if ((f <= ftest || f <= finit + epsf) && std::fabs(g) <= gtol * (-ginit)) return 1;
    if (brackt && (stp <= stmin || stp >= stmax)) return 2;
    if (brackt && stmax - stmin <= xtol * stmax) return 2;
    if (stp == stpmax && f <= ftest && g <= gtest) return 2;
    if (stp == stpmin && (f > ftest || g >= gtest)) return 2;
    if (f <= ftest && std::fabs(g) <= gtol * (-ginit)) return 1;
    if (stage == 1 && f <= fx && f > ftest) {
      double fm = f - stp * gtest, fxm = fx - stx * gtest, fym = fy - sty * gtest;
      double gm = g - gtest, gxm = gx - gtest, gym = gy - gtest;
      cstep(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
      fx = fxm + stx * gtest; fy = fym + sty * gtest;
      gx = gxm + gtest; gy = gym + gtest;
    } else {
      cstep(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
    }
    if (brackt) {
      if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
      width1 = width;
      width = std::fabs(sty - stx);
    }
    if (brackt) {
      stmin = std::min(stx, sty);
      stmax = std::max(stx, sty);
    } else {
      stmin = stp + 1.1 * (stp - stx);
      stmax = stp + 4.0 * (stp - stx);
    }
    stp = std::max(stp, stpmin);
    stp = std::min(stp, stpmax);
    if ((brackt && (stp <= stmin || stp >= stmax)) || (brackt && stmax - stmin <= xtol * stmax)) stp = stx;
    return 0;
  }
};


}  // namespace a2c
