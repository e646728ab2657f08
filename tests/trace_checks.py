"""Closest-hit agreement of an FP32 traversal with the oracle's FP64 World::hit (world.rs:81-121).

Rays come in two sets:

* path rays -- the rays a render traces: camera rays of random pixels, then up to two bounces
  leaving each hit point along a cosine-distributed direction about the oriented normal from
  hit.p + 1e-4 d (material.rs:82-95);
* stress rays -- random origins and directions in a box around the scene (origins may lie inside
  closed meshes, where faces are coplanar with others).

A disagreement is classified as
  tie     -- the oracle's own FP64 distance to the FP32 traversal's primitive is within 4 FP32 ulps
             (2^-21 relative) of its closest distance: FP32 cannot order the two (coplanar faces,
             shared edges);
  edge    -- the ray crosses the plane of the oracle's triangle, or of the FP32 traversal's, within
             1e-5 (barycentric) of one of its edges: FP32 cannot tell whether it hits that triangle,
             so the ray may pass between two neighbours and report what lies behind;
  eps     -- the oracle returns the FP32 answer with tmin scaled by 1/2 .. 16: the reference's instance
             test takes each instance's closest object-space hit and drops the whole instance when
             that hit lies under tmin in world space (world.rs:92-108), so a bounce ray leaving a
             scaled-down instance can skip that instance's farther surfaces;
  nudged  -- the oracle itself returns the FP32 traversal's answer for the ray moved by 1e-5 along
             some axis (grazing rays, silhouettes, the reference's own box tests at an edge);
  other   -- none of these: a traversal error.
"""
import numpy as np

import oracle

TIE_REL = 2.0 ** -21
EDGE_BARY = 1e-5


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def stress_rays(rng, n, lo, hi, targets=None, tmin=1e-4, tmax=1e30):
    o = rng.uniform(lo, hi, (n, 3))
    if targets is None:
        d = unit(rng.standard_normal((n, 3)))
    else:
        t = targets[rng.integers(len(targets), size=n)] + rng.normal(0, 0.3, (n, 3))
        d = unit(t - o)
    r = np.zeros((n, 8))
    r[:, :3], r[:, 3:6], r[:, 6], r[:, 7] = o, d, tmin, tmax
    # the FP32 traversal sees FP32 rays: the oracle gets the same rounded values
    return r.astype(np.float32).astype(np.float64)


def path_rays(sc, rng, n, bounces=2, threads=8, crop=None, orc=None):
    """n camera rays through random points of the image (or of a crop x0, y0, x1, y1 in pixels
    of a w x h image, crop = (x0, y0, x1, y1, w, h)), then `bounces` generations of diffuse
    bounce rays from their hits (each generation as many rays as the previous one hit).  `orc`:
    the scene's OracleScene, if already built."""
    from tinypt_amd import _abi
    c = _abi.CameraDesc()
    c.eye[:], c.target[:], c.up[:] = sc.cam.eye, sc.cam.target, sc.cam.up
    c.vfov_deg, c.aspect, c.focus_dist, c.aperture = sc.cam.vfov_deg, sc.cam.aspect, sc.cam.focus_dist, sc.cam.aperture
    if crop is None:
        sxy = rng.uniform(-1, 1, (n, 2))
    else:
        x0, y0, x1, y1, w, h = crop
        sxy = np.stack([rng.uniform(x0, x1, n) * 2 / w - 1, 1 - rng.uniform(y0, y1, n) * 2 / h], 1)
    cam = oracle.camera_rays(c, int(rng.integers(1 << 30)), sxy)
    r = np.zeros((n, 8))
    r[:, :6], r[:, 6], r[:, 7] = cam[:, :6], 1e-4, 1e30
    gens = [r.astype(np.float32).astype(np.float64)]
    orc = orc if orc is not None else oracle.OracleScene(sc)
    for _ in range(bounces):
        h = orc.trace(gens[-1], threads=threads)
        ok = h["kind"] != 0
        if not ok.any():
            break
        p, nrm, d_in = h["p"][ok], h["n"][ok], gens[-1][ok, 3:6]
        nrm = np.where((np.einsum("ij,ij->i", nrm, d_in) < 0)[:, None], nrm, -nrm)
        # cosine-distributed about the oriented normal
        u1, u2 = rng.random(len(p)), rng.random(len(p))
        a = unit(np.where(np.abs(nrm[:, :1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]]) - 0 * nrm)
        t1 = unit(np.cross(nrm, a))
        t2 = np.cross(nrm, t1)
        rr, ph = np.sqrt(u1), 2 * np.pi * u2
        d = t1 * (rr * np.cos(ph))[:, None] + t2 * (rr * np.sin(ph))[:, None] + nrm * np.sqrt(1 - u1)[:, None]
        d = unit(d)
        q = np.zeros((len(p), 8))
        q[:, :3], q[:, 3:6], q[:, 6], q[:, 7] = p + 1e-4 * d, d, 1e-4, 1e30
        gens.append(q.astype(np.float32).astype(np.float64))
    return np.concatenate(gens)


def classify(sc, rays, got, threads=8, orc=None):
    """got: dict(t, kind, inst, prim) of the FP32 traversal.  Returns a report dict: raw
    agreement, counts per class, and the error of the distance of agreeing hits, relative to
    max(t, 1).  `orc`: the scene's OracleScene, if already built."""
    orc = orc if orc is not None else oracle.OracleScene(sc)
    r = orc.trace(rays, threads=threads)
    same = (got["kind"] == r["kind"]) & (got["inst"] == r["inst"]) & (got["prim"] == r["prim"])
    bad = np.nonzero(~same)[0]
    both = (got["kind"][bad] != 0) & (r["kind"][bad] != 0)
    t_got_prim, m_got = orc.prim_distance(rays[bad], got["kind"][bad], got["inst"][bad], got["prim"][bad])
    _, m_orc = orc.prim_distance(rays[bad], r["kind"][bad], r["inst"][bad], r["prim"][bad])
    tie = both & (np.abs(t_got_prim - r["t"][bad]) <= TIE_REL * r["t"][bad])
    edge = ~tie & ((np.abs(m_orc) < EDGE_BARY) & (r["kind"][bad] == 1) | (np.abs(m_got) < EDGE_BARY) & (got["kind"][bad] == 1))
    eps, nudged, other = [], [], []

    def same_as_got(q, i):
        rq = orc.trace(q, threads=1)
        return bool(rq["kind"][0] == got["kind"][i] and rq["inst"][0] == got["inst"][i] and rq["prim"][0] == got["prim"][i])

    for i in bad[~tie & ~edge]:
        q = rays[i:i + 1].copy()
        if any(same_as_got(np.concatenate([q[:, :6], q[:, 6:7] * f, q[:, 7:8]], 1), i) for f in (0.5, 2.0, 4.0, 8.0, 16.0)):
            eps.append(int(i))
            continue
        ok = False
        for ax in range(3):
            for s in (-1e-5, 1e-5):
                q = rays[i:i + 1].copy()
                q[0, ax] += s * max(1.0, abs(q[0, ax]))
                ok = ok or same_as_got(q, i)
        (nudged if ok else other).append(int(i))
    hit = same & (r["kind"] != 0)
    rel = np.abs(got["t"][hit].astype(np.float64) - r["t"][hit]) / np.maximum(r["t"][hit], 1.0)
    return dict(n=len(rays), hit=float((r["kind"] != 0).mean()), agree=float(same.mean()), ties=int(tie.sum()),
                edge=int(edge.sum()), eps=len(eps), nudged=len(nudged), other=other,
                t_rel_p999=float(np.quantile(rel, 0.999)) if hit.any() else 0.0,
                t_rel_max=float(rel.max()) if hit.any() else 0.0)


def report(name, rep):
    print(f"{name}: {rep['n']} rays, hit {rep['hit']:.3f}, raw agreement {rep['agree']:.6f}; disagreements: "
          f"ties {rep['ties']}, edge {rep['edge']}, eps {rep['eps']}, nudged {rep['nudged']}, other {len(rep['other'])}; "
          f"t rel err p99.9 {rep['t_rel_p999']:.2e} max {rep['t_rel_max']:.2e}", flush=True)
