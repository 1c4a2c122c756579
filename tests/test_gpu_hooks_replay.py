"""The drop-in seam exercised as the reference drives it: piv_pass's hook
sequence (src/pipeline.jl:655-832) replayed through the C ABI, with the host
steps a Hammerhead.jl caller keeps (the Interpolations prefilter, predictor
node values, validate_and_replace!, build_predictor's smoothing) taken from
the oracle's Float32 restatement of the reference and the device steps
through the hooks:

  _deform_context  -> hh_set_coefs (the host interpolants' coefficients, once
                      per run), apply_predictor -> hh_deform per sweep,
  process_windows! -> hh_process_windows (HH_SRC_WARP on the device-resident
                      warp, or a host pair uploaded on every call),
  uncertainty_sweep! -> hh_uncertainty_sweep after the sweep loop.

Checked: the replay equals hh_run_piv's device schedule to 1e-4 px and meets
the parity bars against the FP64 oracle; host-deformed warps in buffers
reused in place for every sweep (the reference's warp_buffers) agree with
the device-resident path to 1e-4 px, because a host pair is uploaded on every
call; two pairs replayed on one context give bitwise the result of each on a
fresh context.
"""
import numpy as np
import pytest

from parity_bars import field_parity, say

pytestmark = pytest.mark.gpu


def schedule(hh):
    return hh.multipass_parameters([64, 32, 16], padding=True, apodization="gauss",
                                   final=dict(overlap=12, max_iterations=3, convergence_tol=0.0, uncertainty=True))


def replay(hh, orc, ctx, A, B, passes, host_warp=False):
    nr, nc = A.shape
    A = np.asfortranarray(A, dtype=np.float32)
    B = np.asfortranarray(B, dtype=np.float32)
    # image_interpolant on the host (Interpolations' prefilter in Float32)
    cA, cB = orc.prefilter(A, prec="f32"), orc.prefilter(B, prec="f32")
    coefs_staged = False
    warpA = np.zeros((nr, nc), np.float32, order="F")  # warp_buffers, reused in place
    warpB = np.zeros((nr, nc), np.float32, order="F")
    pred = None
    res = None
    for k, p in enumerate(passes):
        force_replace = k < len(passes) - 1
        x, y, jobs = hh.pass_grid(p, nr, nc)
        ny, nx = y.size, x.size
        maxiter = p.max_iterations
        nalt = max(p.n_peaks - 1, 1)
        g = lambda: np.zeros((ny, nx), np.float32, order="F")
        out = dict(u=g(), v=g(), peak_ratio=g(), correlation_moment=g(),
                   alt_u=np.full((ny, nx, nalt), np.nan, np.float32, order="F"),
                   alt_v=np.full((ny, nx, nalt), np.nan, np.float32, order="F"))
        unc = p.uncertainty and not force_replace
        fused = unc and maxiter == 1
        uu = np.full((ny, nx), np.nan, np.float32, order="F")
        uv = np.full((ny, nx), np.nan, np.float32, order="F")
        if fused:
            out["unc_u"], out["unc_v"] = uu, uv
        keep_measured = maxiter > 1 and not (force_replace or p.replace_outliers)
        for it in range(1, maxiter + 1):
            if pred is None:
                ctx.set_pair(A, B)  # a host pair: uploaded on every call
                src = hh.HH_SRC_RAW
                u = g()
                v = g()
            else:
                py, px, pu, pv = pred
                u, v = orc.node_values(py, px, pu, pv, y, x, prec="f32")
                if host_warp:
                    # apply_predictor without a device context: the host
                    # deforms into the reused warp buffers, and the pair is
                    # uploaded on this call
                    wA, wB = orc.deform(cA, cB, nr, nc, py, px, pu, pv, prec="f32")
                    warpA[...] = wA
                    warpB[...] = wB
                    ctx.set_pair(warpA, warpB)
                    src = hh.HH_SRC_RAW
                else:
                    if not coefs_staged:
                        ctx.set_coefs(cA, cB)
                        coefs_staged = True
                    ctx.deform(py, px, pu, pv)
                    src = hh.HH_SRC_WARP
            if it > 1:
                for key in ("u", "v"):
                    out[key][...] = 0
                out["alt_u"][...] = np.nan
                out["alt_v"][...] = np.nan
            ctx.process_windows(p, jobs, ny, nx, src=src, uncertainty=fused, out=out)
            alt_u = out["alt_u"] + u[:, :, None]
            alt_v = out["alt_v"] + v[:, :, None]
            u = (u + out["u"]).astype(np.float32)
            v = (v + out["v"]).astype(np.float32)
            if keep_measured:
                mu, mv = u.copy(), v.copy()
            u, v, outl = orc.validate_and_replace(p.to_struct(), u, v, out["peak_ratio"], out["correlation_moment"],
                                                  alt_u if p.n_peaks > 1 else None,
                                                  alt_v if p.n_peaks > 1 else None,
                                                  force_replace=force_replace or maxiter > 1, prec="f32")
            if it == maxiter:
                break
            pred = (y, x, orc.smooth_field(u, prec="f32"), orc.smooth_field(v, prec="f32"))
        if unc and not fused:
            uu, uv = ctx.uncertainty_sweep(p, jobs, ny, nx, src=src)
        if keep_measured:
            u[outl], v[outl] = mu[outl], mv[outl]
        res = dict(x=x, y=y, u=u, v=v, peak_ratio=out["peak_ratio"].copy(),
                   correlation_moment=out["correlation_moment"].copy(), uncertainty_u=uu, uncertainty_v=uv,
                   outliers=outl.astype(bool), mask=np.zeros((ny, nx), bool))
        pred = (y, x, orc.smooth_field(u, prec="f32"), orc.smooth_field(v, prec="f32"))
    return res


def _same(a, b):
    return all(np.array_equal(a[k], b[k], equal_nan=True)
               for k in ("u", "v", "peak_ratio", "correlation_moment", "uncertainty_u", "uncertainty_v", "outliers"))


def test_hook_replay_matches_device_schedule_and_oracle(hh, orc):
    pair = hh.synthetic.config_vortex_1024(n=512)
    passes = schedule(hh)
    with hh.Context(0) as c:
        say("hook replay, device-resident warp")
        r = replay(hh, orc, c, pair.A, pair.B, passes)
        c.set_pair(pair.A, pair.B)
        d = c.run_piv(passes)
    assert np.array_equal(r["outliers"], d["outliers"])
    dev = max(np.abs(r["u"] - d["u"]).max(), np.abs(r["v"] - d["v"]).max())
    say("replay vs hh_run_piv max |d| px", dev)
    assert dev < 1e-4
    o = orc.run_piv(pair.A.astype(np.float64), pair.B.astype(np.float64), [p.to_struct() for p in passes],
                    uq_pert=True)
    field_parity(r, o, label="hook replay", rms_tol=0.005)


def test_host_warp_buffers_reused_in_place(hh, orc):
    # the host-deformed warp in buffers reused in place for every sweep and
    # pass: each call uploads the current content, so the field agrees with
    # the device-resident path (a stale upload would be off by pixels)
    pair = hh.synthetic.config_vortex_1024(n=256)
    passes = schedule(hh)
    with hh.Context(0) as c:
        a = replay(hh, orc, c, pair.A, pair.B, passes)
    with hh.Context(0) as c:
        b = replay(hh, orc, c, pair.A, pair.B, passes, host_warp=True)
    assert np.array_equal(a["outliers"], b["outliers"])
    dev = max(np.abs(a["u"] - b["u"]).max(), np.abs(a["v"] - b["v"]).max())
    say("host warp vs device warp max |d| px", dev)
    assert dev < 1e-4


def test_two_pairs_on_one_context(hh, orc):
    p1 = hh.synthetic.config_vortex_1024(n=256)
    p2 = hh.synthetic.particle_pair(256, 256, hh.synthetic.uniform_flow(-1.4, 2.2), seed=77)
    passes = schedule(hh)
    with hh.Context(0) as c:
        r1 = replay(hh, orc, c, p1.A, p1.B, passes)
        r2 = replay(hh, orc, c, p2.A, p2.B, passes)
    with hh.Context(0) as c:
        f2 = replay(hh, orc, c, p2.A, p2.B, passes)
    assert _same(r2, f2)
    assert not np.array_equal(r1["u"], r2["u"])
