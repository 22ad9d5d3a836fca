"""Parity criteria shared by the GPU tests and __graft_entry__.smoke().

TEST INFRASTRUCTURE ONLY (like the rest of oracle/).
"""
import numpy as np

# Oracle "ill-conditioned trajectory" bits (oracle/ocpx_oracle.c, pe_t.ill):
#   1 stability SSM at its cap, 2 stability NR at its cap, 4 split SSM at its
#   cap, 8 split NR unconverged, 16 non-finite residual, 32 Rachford-Rice
#   fraction outside [-10, 11], 64 CheckSplit fallback, 128 Rachford-Rice at
#   its cap, 256 split restart cap.
# Building the same oracle source with fused multiply-add contraction flips
# phase labels only in cells carrying one of the split / Rachford-Rice bits
# (4 and up); a stability iteration stopped at its cap alone never did.  So
# labels must agree on every cell without those bits.
SPLIT_ILL = ~3


def rounding_sensitive(ill):
    return (np.asarray(ill) & SPLIT_ILL) != 0


# Scale floor of the relative error: an entry is compared relative to
# max(|b|, FLOOR * max|b|), the max taken over its field or its row.  Entries
# more than 1e4 times below their field's (row's) largest entry are sums
# whose terms are of that largest size; rounding leaves them an absolute error
# of a few ulp of that size, so they are held to PARITY_RTOL * FLOOR = 1e-14
# of it (about 50 ulp) instead of to 1e-10 of their own size.  The tests
# count, per field, the entries that pass only because of the floor.
FLOOR = 1e-4


def rel_err(a, b, rowwise=False, floor=FLOOR):
    """|a - b| / (|b| + floor max|b|), the max taken over the whole field or
    over each row; exact equality counts as zero.  floor = 0: plain relative
    error."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if b.size == 0:
        return np.zeros_like(b)
    if rowwise:
        scale = floor * np.max(np.abs(b), axis=1, keepdims=True)
    else:
        scale = floor * np.max(np.abs(b))
    e = np.abs(a - b) / (np.abs(b) + scale + 1e-300)
    e[a == b] = 0
    return e


# Per-entry parity bar: 1e-10 relative, unless the reference algorithm itself
# is not reproducible to that level on the entry -- measured by re-running the
# oracle under rounding-level perturbations (the fused-multiply-add build, and
# the pressure scaled by 1 + 2^-50).  There the bar is SPREAD_MULT times the
# spread the perturbations produce: the device's exp/log round differently
# from the host's, a perturbation of the same kind.
PARITY_RTOL = 1e-10
SPREAD_MULT = 100.0


def rounding_spread(st, others, field, n, rowwise=False, floor=FLOOR):
    ref = st.a[field].reshape(n, -1)
    lab = st.a["PHASEEXIST"].reshape(n, 3)
    spread = np.zeros_like(ref, dtype=float)
    for o in others:
        same = np.all(o.a["PHASEEXIST"].reshape(n, 3) == lab, axis=1)
        e = rel_err(o.a[field].reshape(n, -1), ref, rowwise, floor)
        e[~same] = 0
        spread = np.maximum(spread, e)
    return spread


# Trajectories through degenerate iterates: a Rachford-Rice fraction outside
# [-10, 11] (bit 32) makes the phase compositions z / (1 + nu (K - 1))
# negative or huge, and their fugacities overflow or turn non-finite (bit
# 16).  How such a trajectory continues depends on the order in which each
# operation rounds and overflows: the kernels' own logic stepped on the host
# with fused multiply-adds flips labels in cells that the oracle's
# perturbations leave alone.  Labels there are compared but not required to
# agree; the tests bound how many differ.
DEGENERATE = 16 | 32


def degenerate(ill):
    return (np.asarray(ill) & DEGENERATE) != 0


def cell_spread(spread):
    """The largest entry of each cell's spread: how far rounding moves the
    cell's field as a whole, which its solves' conditioning sets."""
    return np.broadcast_to(np.max(spread, axis=1, keepdims=True), spread.shape)


def label_stable(st, others):
    """Cells whose phase labels the oracle reproduces under every rounding-
    level perturbation in others: the labels the reference algorithm itself
    fixes on these inputs, whatever its trajectory (capped or not)."""
    n = st.nb
    lab = st.a["PHASEEXIST"].reshape(n, 3)
    ok = np.ones(n, bool)
    for o in others:
        ok &= np.all(o.a["PHASEEXIST"].reshape(n, 3) == lab, axis=1)
    return ok


def perturbed_oracle_states(oracle, oracle_fma, fluid, st_pre):
    """Flash outputs of the oracle under rounding-level perturbations of the
    pre-flash state st_pre (left untouched): its fused-multiply-add build, and
    the pressures scaled by 1 + 2^-50 and 1 - 2^-50 (plain and fused builds)."""
    out = []
    for orc, f in ((oracle_fma, 1.0), (oracle, 1 + 2.0 ** -50), (oracle, 1 - 2.0 ** -50),
                   (oracle_fma, 1 - 2.0 ** -50)):
        a = st_pre.copy()
        a.a["P"] *= f
        orc.cal_flash(fluid, a)
        out.append(a)
    return out
