"""Parity criteria between the HIP path and the oracle (the CPU restatement).

* spike rasters identical up to spike times within one dt (one row);
* voltage traces within TOL_MV max-abs wherever the rasters agree: per cell, up to the first
  row at which its raster is not identical (for networks: at which any raster is not
  identical, since every cell may be downstream of it).
"""
import numpy as np

TOL_MV = 1e-3


def rasters_by_col(cols, rows, n_cols):
    out = [[] for _ in range(n_cols)]
    for c, r in zip(cols.tolist(), rows.tolist()):
        out[c].append(r)
    return out


def first_raster_mismatch(ra, rb, slack=1):
    """Row of the first spike that is not matched within `slack` rows, or None."""
    i = j = 0
    while i < len(ra) and j < len(rb):
        if abs(ra[i] - rb[j]) <= slack:
            i += 1
            j += 1
            continue
        return min(ra[i], rb[j])
    if i < len(ra):
        return ra[i]
    if j < len(rb):
        return rb[j]
    return None


def compare(trace_gpu, ev_gpu, trace_ref, ev_ref, n_event_cols, record_to_event_col=None, network=False):
    """Return a dict of parity figures. record_to_event_col maps trace columns to the event
    column of the same cell (None: no spike gating)."""
    rows = min(trace_gpu.shape[0], trace_ref.shape[0])
    ra = rasters_by_col(*ev_gpu, n_event_cols)
    rb = rasters_by_col(*ev_ref, n_event_cols)
    mism = [first_raster_mismatch(a, b) for a, b in zip(ra, rb)]          # beyond one dt
    n_mismatch = sum(m is not None for m in mism)
    exact = [first_raster_mismatch(a, b, slack=0) for a, b in zip(ra, rb)]  # any difference
    global_cut = rows
    if network:
        ms = [m for m in exact if m is not None]
        if ms:
            global_cut = min(ms)
    max_err = 0.0
    compared = 0
    for c in range(trace_gpu.shape[1]):
        cut = global_cut
        if record_to_event_col is not None and not network:
            m = exact[record_to_event_col[c]]
            if m is not None:
                cut = min(cut, m)
        if cut > 0:
            d = np.abs(trace_gpu[:cut, c].astype(np.float64) - trace_ref[:cut, c].astype(np.float64))
            max_err = max(max_err, float(d.max()))
            compared += cut
    exact_spikes = sum(len(a) for a in ra)
    same = sum(1 for a, b in zip(ra, rb) if a == b)
    return dict(rows=rows, max_abs_mV=max_err, compared_samples=compared, raster_mismatch_cols=n_mismatch,
                identical_raster_cols=same, spikes_gpu=exact_spikes, spikes_ref=sum(len(b) for b in rb),
                first_mismatch_row=min([m for m in mism if m is not None], default=None))


def compare_fine(trace, ev, trace_fine, ev_fine, substeps, n_event_cols, record_to_event_col):
    """A trace at dt against a run at dt / substeps that recorded every substeps-th row
    (event rows of the fine run in fine steps): the fine spike rows are taken to the coarse
    row at the same time (rounded up), rasters compared within one coarse dt, voltages
    compared per cell up to its first spike that differs."""
    cols, rows = ev_fine
    coarse = (np.asarray(rows) + substeps - 1) // substeps
    return compare(trace, ev, trace_fine, (np.asarray(cols), coarse), n_event_cols, record_to_event_col)
