"""Shared parity helpers of the GPU tests (test infrastructure): the oracle run on the
device's line-search path, and the oracle's own rounding spread on that path."""
import oracle as o
import ptn_hip as P


def say(*a):
    print("[gpu-test]", *a, flush=True)


def forced_outcomes(logg):
    """The device's line-search outcomes per sweep (-1 where none was tried)."""
    return [l["accel_accepted"] if l["accel_tried"] else -1 for l in logg]


def oracle_on_device_path(X, R, init, logg, tie, **kw):
    """The oracle's cpd() following the device's line-search outcomes; every sweep where
    they differ from the oracle's own test f > f_accel must be a tie to `tie` (relative)."""
    Fo, sto, logo = o.cpd(X, R, init, force_accept=forced_outcomes(logg), **kw)
    for k, l in enumerate(logo):
        if l["accel_overridden"]:
            gap = abs(l["f_value"] - l["f_accel"]) / l["f_value"]
            say("line-search tie at sweep %d: oracle |f - f_accel| / f = %.2e" % (k + 1, gap))
            assert gap <= tie, (k, l["f_value"], l["f_accel"])
    return Fo, sto, logo


def oracle_spread(X, R, init, logg, Fo, logo, oform, **kw):
    """How far the oracle moves from its own result when only the summation order of the
    MTTKRPs changes: the largest factor (relative) and cost (|f^2 - f^2| / ||X||^2) gaps
    between the reference run (Fo, logo; form oform) and the oracle's other deterministic
    forms (per-mode, single-thread blocked, dimension tree for orders 3 and 4), all on the
    device's line-search path.  This is the rounding sensitivity of the problem itself."""
    N = len(X.shape)
    forms = [(o.MTTKRP_STANDARD, 0), (o.MTTKRP_BLOCKED, 1)] + ([(o.MTTKRP_DIMTREE, 0)] if N in (3, 4) else [])
    frob = float((X * X).sum())
    fs = cs = 0.0
    for form, threads in forms:
        if form == oform:
            continue
        Fa, _, loga = o.cpd(X, R, init, force_accept=forced_outcomes(logg), mttkrp_form=form, threads=threads, **kw)
        fs = max(fs, max(P.rel_err(a, b) for a, b in zip(Fa, Fo)))
        cs = max(cs, max(abs(a["f_value"] ** 2 - b["f_value"] ** 2) for a, b in zip(loga, logo)) / frob)
    return fs, cs
