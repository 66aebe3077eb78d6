/*
 * edenhip_setup.h -- host-side setup helpers exported by libedenhip.so.
 *
 * These are the parts of EDEN's model builder that define the numbers the per-step loop
 * consumes and that a caller without EDEN's GenerateModel (the standalone CLI, the Python
 * package) needs in order to fill an edenhip_model_desc.
 */
#ifndef EDENHIP_SETUP_H
#define EDENHIP_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Cable constants of one cell whose morphology has one compartment per segment
 * (free-standing segments, eden/Eden.cpp:1437-1491). Replaces the geometry and cable
 * analysis of eden/Eden.cpp:4015-4501 (areas, capacitances, midpoint-split axial
 * resistances, per-compartment sum of 1/RC over neighbours).
 *   n              segments (= compartments), parents before children, parent[0] = -1
 *   proximal/distal [n*4] x, y, z, diameter in um (the caller resolves NeuroML's implicit
 *                  proximal points)
 *   cm             [n] specific capacitance, uF/cm^2
 *   ra             [n] axial resistivity, kOhm*cm
 * Outputs (each may be NULL): area um^2, capacitance pF, resistance to parent MOhm (NaN for
 * the root), 1/RC 1/us. Returns 0, or EDENHIP_E_INVALID for a malformed tree or an
 * undefined inter-compartment resistance.
 */
int edenhip_cable_setup(int32_t n, const int32_t *parent, const float *proximal, const float *distal,
                        const float *cm, const float *ra, float *area_out, float *cap_out,
                        float *r_axial_out, float *inv_rc_out);

/*
 * Cable constants of one cell whose compartments are made of segment slices: EDEN's cable
 * discretisation (eden/Eden.cpp:1387-1772) makes each cable segment group (neuroLexId
 * sao864921383) nseg compartments of equal path length, each spanning parts of one or more
 * segments, and every other segment one compartment. Replaces eden/Eden.cpp:4015-4501 for
 * such compartments: lengths and areas summed over the slices, capacitance from each slice's
 * segment Cm, the proximal / distal half resistances split at the compartment's path-length
 * midpoint, R to parent = R_proximal(child) + R_distal(parent), 1/RC over [parent, children
 * ascending].
 *   n_seg, proximal/distal [n_seg*4], cm [n_seg], ra [n_seg]   the segments, as above
 *   n_comp, parent [n_comp]      compartments, parents before children, parent[0] = -1
 *   slice_begin [n_comp + 1]     each compartment's slices, in path order from its proximal end:
 *   slice_seg, slice_from, slice_to   segment and fractionAlong range [from, to] of each slice
 * Outputs per compartment as edenhip_cable_setup (each may be NULL).
 */
int edenhip_cable_setup_slices(int32_t n_seg, const float *proximal, const float *distal, const float *cm,
                               const float *ra, int32_t n_comp, const int32_t *parent, const int32_t *slice_begin,
                               const int32_t *slice_seg, const float *slice_from, const float *slice_to,
                               float *area_out, float *cap_out, float *r_axial_out, float *inv_rc_out);

#ifdef __cplusplus
}
#endif
#endif /* EDENHIP_SETUP_H */
