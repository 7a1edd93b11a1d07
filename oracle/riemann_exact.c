/*
 * TEST INFRASTRUCTURE (the oracle): Toro's exact solution of the 1D ideal-gas Riemann problem,
 * restated from src/reference/riemann_exact.ts:9-71 (Newton iteration for p* from the
 * two-rarefaction guess, then sampling of the self-similar fan). Used to check shock tubes.
 *
 * Parity pins (tests/test_oracle.py): Toro's table 4.2 Sod star state p* = 0.30313,
 * u* = 0.92745, rho = 0.26557 between contact and shock; Toro test 3 p* = 460.894,
 * u* = 19.5975 (tests/unit/reference.test.ts:20-35); and the Rankine-Hugoniot jump conditions,
 * checked across the sampled shock.
 */
#include <math.h>
#include "oracle.h"

static void fk(double p, const double s[3], double c, double g, double out[2]) {
  if (p > s[2]) {
    double A = 2 / ((g + 1) * s[0]);
    double B = ((g - 1) / (g + 1)) * s[2];
    double q = sqrt(A / (p + B));
    out[0] = (p - s[2]) * q;
    out[1] = q * (1 - (p - s[2]) / (2 * (B + p)));
    return;
  }
  double pr = p / s[2];
  out[0] = ((2 * c) / (g - 1)) * (pow(pr, (g - 1) / (2 * g)) - 1);
  out[1] = (1 / (s[0] * c)) * pow(pr, -(g + 1) / (2 * g));
}

void orc_exact_riemann(const double L[3], const double R[3], double g, double xi, double out[3]) {
  double cL = sqrt((g * L[2]) / L[0]);
  double cR = sqrt((g * R[2]) / R[0]);
  double z = (g - 1) / (2 * g);
  double p = pow((cL + cR - 0.5 * (g - 1) * (R[1] - L[1])) / (cL / pow(L[2], z) + cR / pow(R[2], z)), 1 / z);
  if (p < 1e-12) p = 1e-12;
  for (int i = 0; i < 100; i++) {
    double fl[2], fr[2];
    fk(p, L, cL, g, fl);
    fk(p, R, cR, g, fr);
    double pn = p - (fl[0] + fr[0] + R[1] - L[1]) / (fl[1] + fr[1]);
    if (pn < 1e-14) pn = 1e-14;
    double change = (2 * fabs(pn - p)) / (pn + p);
    p = pn;
    if (change < 1e-14) break;
  }
  double ps = p;
  double a[2], b[2];
  fk(ps, R, cR, g, a);
  fk(ps, L, cL, g, b);
  double us = 0.5 * (L[1] + R[1]) + 0.5 * (a[0] - b[0]);
  double gm = (g - 1) / (g + 1);
  if (xi <= us) {
    if (ps > L[2]) {
      double sL = L[1] - cL * sqrt(((g + 1) / (2 * g)) * (ps / L[2]) + (g - 1) / (2 * g));
      if (xi <= sL) { out[0] = L[0]; out[1] = L[1]; out[2] = L[2]; return; }
      out[0] = (L[0] * (ps / L[2] + gm)) / (gm * (ps / L[2]) + 1); out[1] = us; out[2] = ps;
      return;
    }
    double head = L[1] - cL;
    double cs = cL * pow(ps / L[2], z);
    double tail = us - cs;
    if (xi <= head) { out[0] = L[0]; out[1] = L[1]; out[2] = L[2]; return; }
    if (xi >= tail) { out[0] = L[0] * pow(ps / L[2], 1 / g); out[1] = us; out[2] = ps; return; }
    double c = (2 / (g + 1)) * (cL + 0.5 * (g - 1) * (L[1] - xi));
    out[0] = L[0] * pow(c / cL, 2 / (g - 1));
    out[1] = (2 / (g + 1)) * (cL + 0.5 * (g - 1) * L[1] + xi);
    out[2] = L[2] * pow(c / cL, (2 * g) / (g - 1));
    return;
  }
  if (ps > R[2]) {
    double sR = R[1] + cR * sqrt(((g + 1) / (2 * g)) * (ps / R[2]) + (g - 1) / (2 * g));
    if (xi >= sR) { out[0] = R[0]; out[1] = R[1]; out[2] = R[2]; return; }
    out[0] = (R[0] * (ps / R[2] + gm)) / (gm * (ps / R[2]) + 1); out[1] = us; out[2] = ps;
    return;
  }
  double head = R[1] + cR;
  double cs = cR * pow(ps / R[2], z);
  double tail = us + cs;
  if (xi >= head) { out[0] = R[0]; out[1] = R[1]; out[2] = R[2]; return; }
  if (xi <= tail) { out[0] = R[0] * pow(ps / R[2], 1 / g); out[1] = us; out[2] = ps; return; }
  double c = (2 / (g + 1)) * (cR - 0.5 * (g - 1) * (R[1] - xi));
  out[0] = R[0] * pow(c / cR, 2 / (g - 1));
  out[1] = (2 / (g + 1)) * (-cR + 0.5 * (g - 1) * R[1] + xi);
  out[2] = R[2] * pow(c / cR, (2 * g) / (g - 1));
}
