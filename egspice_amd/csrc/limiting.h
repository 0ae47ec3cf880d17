// SPICE limiting shared by the hand-written kinds (devices.h) and the
// compiled ones (va_models.h): ngspice's DEVfetlim, DEVlimvds and DEVpnjlim
// (devsup.c) on one junction or leg, and the snap of a limited image back
// onto the unknowns within 16 ulps of their scale (src/device/eval.zig:
// 1111-1125).
#pragma once

#include "emath.h"

namespace esp {

ESP_HD double fetlim(double vnew, double vold, double vto)
{
    double vtsthi = fabs_(2 * (vold - vto)) + 2;
    double vtstlo = vtsthi / 2 + 2;
    double vtox = vto + 3.5;
    double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) {
                    if (-delv > vtstlo) vnew = vold - vtstlo;
                } else {
                    vnew = va_max(vnew, vto + 2);
                }
            } else {
                if (delv >= vtsthi) vnew = vold + vtsthi;
            }
        } else {
            if (delv <= 0) vnew = va_max(vnew, vto - 0.5);
            else vnew = va_min(vnew, vto + 4);
        }
    } else {
        if (delv <= 0) {
            if (-delv > vtsthi) vnew = vold - vtsthi;
        } else {
            double vtemp = vto + 0.5;
            if (vnew <= vtemp) {
                if (delv > vtstlo) vnew = vold + vtstlo;
            } else {
                vnew = vtemp;
            }
        }
    }
    return vnew;
}

ESP_HD double limvds(double vnew, double vold)
{
    if (vold >= 3.5) {
        if (vnew > vold) vnew = va_min(vnew, 3 * vold + 2);
        else if (vnew < 3.5) vnew = va_max(vnew, 2);
    } else {
        if (vnew > vold) vnew = va_min(vnew, 4);
        else vnew = va_max(vnew, -0.5);
    }
    return vnew;
}

ESP_HD double pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck)
{
    if (vnew > vcrit && fabs_(vnew - vold) > vt + vt) {
        if (vold > 0) {
            double arg = 1 + (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * e_log(arg);
            else vnew = vcrit;
        } else {
            vnew = vt * e_log(vnew / vt);
        }
        *icheck = 1;
    } else {
        if (vnew < 0) {
            double arg = vold > 0 ? -1 * vold - 1 : 2 * vold - 1;
            if (vnew < arg) {
                vnew = arg;
                *icheck = 1;
            } else {
                *icheck = 0;
            }
        } else {
            *icheck = 0;
        }
    }
    return vnew;
}

ESP_HD void snap_limited(int nu, unsigned reads, unsigned writes, double *lx, const double *cur)
{
    double scale = 0;
    for (int u = 0; u < nu; u++)
        if (reads >> u & 1) scale = fmax_(scale, fabs_(cur[u]));
    double tol = 16 * 2.220446049250313e-16 * scale;
    for (int u = 0; u < nu; u++)
        if ((writes & reads) >> u & 1)
            if (fabs_(lx[u] - cur[u]) <= tol) lx[u] = cur[u];
}

}  // namespace esp
