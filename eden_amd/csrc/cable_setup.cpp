// cable_setup.cpp -- morphology -> per-compartment cable constants, host side.
//
// Restates EDEN's discretisation for free-standing segments (one compartment per segment,
// eden/Eden.cpp:1437-1491) and its cable analysis (eden/Eden.cpp:4015-4501):
//   area      frustum lateral surface, or pi d^2 for a zero-length segment (eden/Eden.cpp:237-262)
//   C         Cm * area, (uF/cm^2 * um^2) -> pF                          (eden/Eden.cpp:4222-4227)
//   R         frustum resistance Ra (4/pi) L / (d1 d2), (kOhm cm um/um^2) -> MOhm, split at the
//             compartment midpoint; R to parent = R_proximal(child) + R_distal(parent)
//                                                                          (eden/Eden.cpp:4159-4379)
//   1/RC      per compartment, summed over the adjacency list [parent, children ascending]
//                                                                          (eden/Eden.cpp:4488-4501)
// Floating-point types follow the reference (`Real` = float, M_PI and unit factors in double).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/edenhip.h"
#include "../../include/edenhip_setup.h"

namespace {

struct Pt {
    float x, y, z, d;
};

float lerp(float a, float b, float r) { return a + (b - a) * r; }

Pt lerp_pt(const Pt &a, const Pt &b, float r) { return Pt{lerp(a.x, b.x, r), lerp(a.y, b.y, r), lerp(a.z, b.z, r), lerp(a.d, b.d, r)}; }

float length3(float dx, float dy, float dz) { return std::sqrt(dx * dx + dy * dy + dz * dz); }

float distance(const Pt &a, const Pt &b) { return length3(b.x - a.x, b.y - a.y, b.z - a.z); }

float frustum_area(float length, float dp, float dd) {
    if (length == 0) return (float)(M_PI * dd * dd);
    return (float)((M_PI / 2.0) * (dp + dd) * std::sqrt(((dp - dd) * (dp - dd) / 4.0) + length * length));
}

// Ra in kOhm*cm, points in um -> MOhm: ((kOhm cm) um / um^2) = 10 MOhm
float frustum_resistance(float Ra, const Pt &from, const Pt &to) {
    float resistance = (float)(Ra * (4 / M_PI) * distance(from, to) / (from.d * to.d));
    resistance = (float)(resistance * (1.0 / 1.0) * std::pow(10.0, 7 - 6));
    return resistance;
}

}  // namespace

extern "C" int edenhip_cable_setup(int32_t n, const int32_t *parent, const float *proximal, const float *distal,
                                   const float *cm, const float *ra, float *area_out, float *cap_out,
                                   float *r_axial_out, float *inv_rc_out) {
    if (n <= 0 || !parent || !proximal || !distal || !cm || !ra) return EDENHIP_E_INVALID;
    if (parent[0] != -1) return EDENHIP_E_INVALID;
    for (int32_t i = 1; i < n; i++)
        if (parent[i] < 0 || parent[i] >= i) return EDENHIP_E_INVALID;
    std::vector<float> len(n), area(n), cap(n), rprox(n), rdist(n), rpar(n, NAN);
    for (int32_t i = 0; i < n; i++) {
        const Pt p{proximal[4 * i], proximal[4 * i + 1], proximal[4 * i + 2], proximal[4 * i + 3]};
        const Pt q{distal[4 * i], distal[4 * i + 1], distal[4 * i + 2], distal[4 * i + 3]};
        // one fragment per compartment: fractionAlong 0 -> 1
        const Pt from = lerp_pt(p, q, 0.0f), to = lerp_pt(p, q, 1.0f);
        const float flen = distance(from, to);
        len[i] = 0 + flen;
        const float farea = frustum_area(flen, from.d, to.d);
        area[i] = 0 + farea;
        // (uF/cm^2 * um^2) -> pF: 10^(-2-12) / 10^-12
        float fcap = cm[i] * farea;
        fcap = (float)(fcap * (1.0 / 1.0) * std::pow(10.0, -14 + 12));
        cap[i] = 0 + fcap;
        const float comp_mid = len[i] / 2;
        float running = 0;
        rprox[i] = 0;
        rdist[i] = 0;
        if (len[i] == 0) {
            Pt ext = from;
            ext.y += to.d;
            const Pt mid = lerp_pt(p, ext, 0.5f);
            rprox[i] += frustum_resistance(ra[i], from, mid);
            rdist[i] += frustum_resistance(ra[i], mid, ext);
        } else if (running + flen < comp_mid) {
            rprox[i] += frustum_resistance(ra[i], from, to);
        } else if (comp_mid <= running) {
            rdist[i] += frustum_resistance(ra[i], from, to);
        } else {
            float frac = (comp_mid - running) / flen;
            if (!(frac < 1)) frac = 1;
            if (!(0 < frac)) frac = 0;
            const Pt mid = lerp_pt(from, to, frac);
            rprox[i] += frustum_resistance(ra[i], from, mid);
            rdist[i] += frustum_resistance(ra[i], mid, to);
        }
        if (parent[i] >= 0) rpar[i] = rprox[i] + rdist[parent[i]];
    }
    for (int32_t i = 1; i < n; i++)
        if (!(std::isfinite(rpar[i]) && rpar[i] > 0)) return EDENHIP_E_INVALID;
    // adjacency lists [parent, children ascending]
    std::vector<std::vector<int32_t>> conn(n);
    for (int32_t i = 0; i < n; i++)
        if (parent[i] >= 0) {
            conn[i].push_back(parent[i]);
            conn[parent[i]].push_back(i);
        }
    for (int32_t i = 0; i < n; i++) {
        float inv = 0;
        for (int32_t adj : conn[i]) {
            const int32_t idx = adj > i ? adj : i;
            const float R = rpar[idx];
            const float C = cap[i];
            const double D = (double)(1 / (R * C)) * (1.0 / 1.0) * std::pow(10.0, 0);
            inv = (float)(inv + D);
        }
        if (area_out) area_out[i] = area[i];
        if (cap_out) cap_out[i] = cap[i];
        if (r_axial_out) r_axial_out[i] = rpar[i];
        if (inv_rc_out) inv_rc_out[i] = inv;
    }
    return EDENHIP_OK;
}
