// cable_setup.cpp -- morphology -> per-compartment cable constants, host side.
//
// Restates EDEN's cable analysis (eden/Eden.cpp:4015-4501) for compartments made of segment
// slices (the discretisation of eden/Eden.cpp:1387-1772: a cable group's nseg compartments of
// equal path length, or one compartment per free-standing segment):
//   length,   summed over the compartment's slices, each slice the frustum between two points
//   area      lerped along its segment; lateral surface, or pi d^2 for zero length
//             (eden/Eden.cpp:237-262, 4082-4117)
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

extern "C" int edenhip_cable_setup_slices(int32_t n_seg, const float *proximal, const float *distal, const float *cm,
                                          const float *ra, int32_t n_comp, const int32_t *parent,
                                          const int32_t *slice_begin, const int32_t *slice_seg, const float *slice_from,
                                          const float *slice_to, float *area_out, float *cap_out, float *r_axial_out,
                                          float *inv_rc_out) {
    if (n_seg <= 0 || n_comp <= 0 || !proximal || !distal || !cm || !ra || !parent || !slice_begin || !slice_seg ||
        !slice_from || !slice_to)
        return EDENHIP_E_INVALID;
    if (parent[0] != -1) return EDENHIP_E_INVALID;
    for (int32_t i = 1; i < n_comp; i++)
        if (parent[i] < 0 || parent[i] >= i) return EDENHIP_E_INVALID;
    for (int32_t i = 0; i < n_comp; i++)
        if (slice_begin[i + 1] <= slice_begin[i]) return EDENHIP_E_INVALID;
    for (int32_t k = 0; k < slice_begin[n_comp]; k++)
        if (slice_seg[k] < 0 || slice_seg[k] >= n_seg || !(0 <= slice_from[k] && slice_from[k] <= slice_to[k] &&
                                                           slice_to[k] <= 1))
            return EDENHIP_E_INVALID;
    auto seg_pt = [&](const float *a, int32_t s) { return Pt{a[4 * s], a[4 * s + 1], a[4 * s + 2], a[4 * s + 3]}; };
    std::vector<float> len(n_comp), area(n_comp), cap(n_comp), rprox(n_comp), rdist(n_comp), rpar(n_comp, NAN);
    // lengths and areas, summed over the compartment's segment slices (eden/Eden.cpp:4082-4117)
    for (int32_t i = 0; i < n_comp; i++) {
        len[i] = 0;
        area[i] = 0;
        for (int32_t k = slice_begin[i]; k < slice_begin[i + 1]; k++) {
            const Pt p = seg_pt(proximal, slice_seg[k]), q = seg_pt(distal, slice_seg[k]);
            const Pt from = lerp_pt(p, q, slice_from[k]), to = lerp_pt(p, q, slice_to[k]);
            const float flen = distance(from, to);
            len[i] += flen;
            area[i] += frustum_area(flen, from.d, to.d);
        }
    }
    // capacitance and the resistances of the compartment's proximal and distal halves
    // (eden/Eden.cpp:4159-4379)
    for (int32_t i = 0; i < n_comp; i++) {
        const float comp_mid = len[i] / 2;
        float running = 0;
        cap[i] = 0;
        rprox[i] = 0;
        rdist[i] = 0;
        for (int32_t k = slice_begin[i]; k < slice_begin[i + 1]; k++) {
            const int32_t sg = slice_seg[k];
            const Pt p = seg_pt(proximal, sg), q = seg_pt(distal, sg);
            const Pt from = lerp_pt(p, q, slice_from[k]), to = lerp_pt(p, q, slice_to[k]);
            const float flen = distance(from, to);
            const float farea = frustum_area(flen, from.d, to.d);
            // (uF/cm^2 * um^2) -> pF: 10^(-2-12) / 10^-12
            float fcap = cm[sg] * farea;
            fcap = (float)(fcap * (1.0 / 1.0) * std::pow(10.0, -14 + 12));
            cap[i] += fcap;
            if (len[i] == 0) {
                // zero-length compartment: a cylinder as long as it is wide
                Pt ext = from;
                ext.y += to.d;
                const Pt mid = lerp_pt(p, ext, 0.5f);
                rprox[i] += frustum_resistance(ra[sg], from, mid);
                rdist[i] += frustum_resistance(ra[sg], mid, ext);
            } else if (running + flen < comp_mid) {
                rprox[i] += frustum_resistance(ra[sg], from, to);
            } else if (comp_mid <= running) {
                rdist[i] += frustum_resistance(ra[sg], from, to);
            } else {
                float frac = (comp_mid - running) / flen;
                if (!(frac < 1)) frac = 1;
                if (!(0 < frac)) frac = 0;
                const Pt mid = lerp_pt(from, to, frac);
                rprox[i] += frustum_resistance(ra[sg], from, mid);
                rdist[i] += frustum_resistance(ra[sg], mid, to);
            }
            running += flen;
        }
        if (parent[i] >= 0) rpar[i] = rprox[i] + rdist[parent[i]];
    }
    for (int32_t i = 1; i < n_comp; i++)
        if (!(std::isfinite(rpar[i]) && rpar[i] > 0)) return EDENHIP_E_INVALID;
    // adjacency lists [parent, children ascending]
    std::vector<std::vector<int32_t>> conn(n_comp);
    for (int32_t i = 0; i < n_comp; i++)
        if (parent[i] >= 0) {
            conn[i].push_back(parent[i]);
            conn[parent[i]].push_back(i);
        }
    for (int32_t i = 0; i < n_comp; i++) {
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

// one compartment per segment: each compartment one whole-segment slice
extern "C" int edenhip_cable_setup(int32_t n, const int32_t *parent, const float *proximal, const float *distal,
                                   const float *cm, const float *ra, float *area_out, float *cap_out,
                                   float *r_axial_out, float *inv_rc_out) {
    if (n <= 0 || !parent || !proximal || !distal || !cm || !ra) return EDENHIP_E_INVALID;
    std::vector<int32_t> begin((size_t)n + 1), seg((size_t)n);
    std::vector<float> from((size_t)n, 0.0f), to((size_t)n, 1.0f);
    for (int32_t i = 0; i <= n; i++) begin[(size_t)i] = i;
    for (int32_t i = 0; i < n; i++) seg[(size_t)i] = i;
    return edenhip_cable_setup_slices(n, proximal, distal, cm, ra, n, parent, begin.data(), seg.data(), from.data(),
                                      to.data(), area_out, cap_out, r_axial_out, inv_rc_out);
}
