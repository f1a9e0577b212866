"""Pins the oracle's restatement of the reference adjoint (orc_adjoint_reference_f, AdjointIntersectRasterization.glsl
:111-362 with BackwardCommon.glsl:107-251) on rays that cross one tet, where each of its parts is the derivative
of a well-defined function and FP64 central differences can check it:

  * colours (RGB and density): the derivative of the fragment's output with sub-segment 0 held at its original
    value (the loop s = 3..1 of :204 never differentiates it), the hit points and segment length fixed;
  * positions: the sum of two derivatives.  (i) The segment length |P0 - P1| with each hit point moving with the
    face's vertices at fixed barycentric coordinates (segmentLengthAdjoint's direct term, :139-181), hit colours
    fixed.  (ii) The barycentric coordinates of each hit point held fixed in space moving with the vertices
    (baryAdjoint, :183-251), which changes both the hit colours and the segment length through D(u, v).
    Sub-segment 0 is held fixed in both.

The single-fragment forward below restates the oracle's frag_eval in numpy (checked against oracle.forward in
FP64 first).  What single fragments leave open is the state chaining between fragments (the reference inverts
sub-segments 3..1 only, so the state it passes to the fragment in front still holds the sub-segment 0 of every
fragment behind it): the second test pins it on rays through two to four tets, through the closed form that
chaining gives each tet's gradient (see the comment above it)."""
import numpy as np

import oracle
from difftetvr_amd import meshgen
from tests.helpers import FOVY, ZFAR, ZNEAR, view_from

FT = [(0, 1, 2), (1, 0, 3), (0, 2, 3), (2, 1, 3)]  # face vertices, Common/TetFaceTable.glsl:29-34


def _bary(p0, p1, p2, P):
    area = max(np.linalg.norm(np.cross(p2 - p0, p2 - p1)), 1e-5)
    return np.linalg.norm(np.cross(p2 - p1, P - p1)) / area, np.linalg.norm(np.cross(P - p0, p2 - p0)) / area


def _composite(tseg, col0, col1, att, seg0=None):
    """4 sub-segments front to back; seg0 replaces sub-segment 0's contribution when given."""
    rc = np.zeros(4)
    first = None
    for s in range(4):
        fb, fm, fe = s * 0.25, (s + 0.5) * 0.25, (s + 1.0) * 0.25
        c0 = col0[:3] * (1 - fb) + col1[:3] * fb
        c1 = col0[:3] * (1 - fe) + col1[:3] * fe
        al = col0[3] * (1 - fm) + col1[3] * fm
        cur = oracle.accumulate_linear_const(tseg, c0, c1, al * att)
        if s == 0:
            first = cur
            if seg0 is not None:
                cur = seg0
        rc = rc + (1.0 - rc[3]) * cur
    return rc, first


class Ray:
    """One ray through one tet: faces, ray parameters and barycentrics at the original geometry."""

    def __init__(self, ro, rd, P, C, att):
        self.ro, self.rd, self.P0, self.C0, self.att = ro, rd, P.copy(), C.copy(), att
        self.hit, self.f0, self.f1, t0, t1 = oracle.ray_tet(ro, rd, P)
        if not self.hit:
            return
        self.t0, self.t1 = t0, t1
        self.H = [ro + t0 * rd, ro + t1 * rd]
        self.faces = [FT[self.f0], FT[self.f1]]
        self.lam = [self._lamk(P, k, self.H[k]) for k in range(2)]
        out, self.seg0 = self.forward_full(P, C)
        self.out = out

    @staticmethod
    def _lam(P, face, Hk):
        a, b, c = face
        u, v = _bary(P[a], P[b], P[c], Hk)
        return np.array([u, v, 1.0 - u - v])

    def _lamk(self, P, k, Hk):
        return self._lam(P, self.faces[k], Hk)

    def forward_full(self, P, C):
        """The oracle's frag_eval (unfrozen), used to check this restatement."""
        hit, f0, f1, t0, t1 = oracle.ray_tet(self.ro, self.rd, P)
        H = [self.ro + t0 * self.rd, self.ro + t1 * self.rd]
        cols = [self._lam(P, FT[f], H[k]) @ C[list(FT[f])] for k, f in enumerate((f0, f1))]
        return _composite((t1 - t0) / 4.0, cols[0], cols[1], self.att)

    def colours_frozen0(self, C):
        cols = [self.lam[k] @ C[list(self.faces[k])] for k in range(2)]
        return _composite((self.t1 - self.t0) / 4.0, cols[0], cols[1], self.att, seg0=self.seg0)[0]

    def positions_i(self, P):
        pts = [self.lam[k] @ P[list(self.faces[k])] for k in range(2)]
        cols = [self.lam[k] @ self.C0[list(self.faces[k])] for k in range(2)]
        return _composite(np.linalg.norm(pts[0] - pts[1]) / 4.0, cols[0], cols[1], self.att, seg0=self.seg0)[0]

    def positions_ii(self, P):
        lam = [self._lamk(P, k, self.H[k]) for k in range(2)]
        pts = [lam[k] @ self.P0[list(self.faces[k])] for k in range(2)]
        cols = [lam[k] @ self.C0[list(self.faces[k])] for k in range(2)]
        return _composite(np.linalg.norm(pts[0] - pts[1]) / 4.0, cols[0], cols[1], self.att, seg0=self.seg0)[0]


def _fd(fn, x0, h):
    g = np.zeros_like(x0)
    for idx in np.ndindex(*x0.shape):
        xp, xm = x0.copy(), x0.copy()
        xp[idx] += h
        xm[idx] -= h
        g[idx] = (fn(xp) - fn(xm)) / (2 * h)
    return g


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_reference_adjoint_single_fragment_rays():
    pos32 = np.array([[-0.22, -0.18, 0.15], [0.21, -0.2, 0.12], [0.02, 0.27, 0.03], [0.05, -0.05, -0.24]], np.float32)
    cells = np.array([[0, 1, 2, 3]], np.int32)
    col32 = np.array([[0.9, 0.2, 0.3, 3.0], [0.1, 0.8, 0.4, 1.5], [0.3, 0.3, 0.9, 4.0], [0.7, 0.6, 0.1, 2.2]],
                     np.float32)
    W = H = 28
    eye = (0.45, 0.35, 0.62)
    att = 1.0
    uf, _ = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    fr = oracle.Fragments(cells, pos32, uf, ZNEAR, W, H)
    img, _, cnt = oracle.forward(fr, pos32, col32, uf, att)
    P = pos32.astype(np.float64)
    C = col32.astype(np.float64)
    ro = uf[64:67].astype(np.float64)
    rng = np.random.default_rng(4)
    dimg = np.zeros((H, W, 4), np.float32)
    rays = []
    for y in range(H):
        for x in range(W):
            if cnt[y, x] != 1:
                continue
            rd = oracle.ray_dir(uf, W, H, x, y).astype(np.float64)
            ray = Ray(ro, rd, P, C, att)
            # keep rays well inside the faces: the same faces a tiny turn of the ray away, and hit points away
            # from the face edges, where the unsigned-area barycentrics |cross(.)| have their kink (no derivative)
            if not ray.hit or any(oracle.ray_tet(ro, rd + d, P)[1:3] != (ray.f0, ray.f1)
                                  for d in (np.array([1e-4, 0, 0]), np.array([0, 1e-4, 0]), np.array([0, 0, 1e-4]))):
                continue
            if min(ray.lam[0].min(), ray.lam[1].min()) < 0.02:
                continue
            dimg[y, x] = rng.normal(size=4)
            rays.append((ray, dimg[y, x].astype(np.float64), y, x))
    assert len(rays) > 60
    # the numpy restatement of the forward is the oracle's (FP64; the rays here are the FP32 oracle's, rounded
    # once, hence 1e-6 rather than rounding level)
    img64, _, _ = oracle.forward(fr, P, C, uf.astype(np.float64), att, dtype=np.float64)
    for ray, _, y, x in rays:
        assert np.abs(ray.out - img64[y, x]).max() < 1e-6
    rays = [(r, d) for r, d, _, _ in rays]

    gpos, gcol = oracle.adjoint_reference(fr, pos32, col32, uf, att, img, dimg)

    def loss_c(Cx):
        return sum(float(d @ r.colours_frozen0(Cx)) for r, d in rays)

    def loss_p(Px):
        return sum(float(d @ (r.positions_i(Px) + r.positions_ii(Px))) for r, d in rays)

    fd_col = _fd(loss_c, C, 1e-6)
    fd_pos = _fd(loss_p, P, 1e-7)
    # loss_p adds the two outputs, so its derivative is the sum of the two parts' derivatives
    print(f"reference adjoint vs FD, {len(rays)} rays: colours rel {_rel(gcol, fd_col):.2e}, "
          f"positions rel {_rel(gpos, fd_pos):.2e}", flush=True)
    assert _rel(gcol, fd_col) <= 1e-5    # measured 5e-7 (FP32 oracle against FP64 differences)
    assert _rel(gpos, fd_pos) <= 1e-4    # measured 3e-6
    # and the parts are really what differ from the exact adjoint
    ex_pos, ex_col = oracle.adjoint_exact(fr, pos32, col32, uf, att, dimg)
    assert _rel(gcol, ex_col) > 1e-2


# ---- state chaining between fragments, on rays through two to four tets ------------------------------------
#
# Along one ray, with tets k = 0..n-1 front to back and their sub-segments s_kj (j = 0..3, premultiplied RGBA),
# the reference starts from the final composite and, per tet from the back, inverts sub-segments 3, 2, 1 only.
# Inverting 'over' is exact algebra, so the state it hands to tet k is the true composite of tets 0..k followed by
# sub-segment 0 of every tet behind k; and the image adjoint it carries there has passed back through sub-segments
# 1..3 of the tets behind only.  For tet k's sub-segments 1..3 this makes the reference's gradient the derivative
# of a closed-form function of tet k's parameters:
#     dL . ( Y_k  over  s_k1 over s_k2 over s_k3  over  Q_k ),
#     Y_k frozen, of transmittance  T_k (1 - a(s_k0)) prod_{i > k} (1 - a(s_i0))   (its RGB does not enter),
#     Q_k frozen = s_{k+1,1} over s_{k+1,2} over s_{k+1,3} over s_{k+2,1} over ... (sub-segments 0 left out),
# with T_k the true transmittance in front of tet k.  Summed over tets and rays, FP64 central differences of it
# pin the chaining; the same sum with the damping prod (1 - a(s_i0)) left out, or with the s_i0 kept in Q_k
# (a reference that inverted every sub-segment), must not match.

def _over(a, b):
    return a + (1.0 - a[3]) * b


def _subsegs(tseg, col0, col1, att):
    out = []
    for s in range(4):
        fb, fm, fe = s * 0.25, (s + 0.5) * 0.25, (s + 1.0) * 0.25
        c0 = col0[:3] * (1 - fb) + col1[:3] * fb
        c1 = col0[:3] * (1 - fe) + col1[:3] * fe
        al = col0[3] * (1 - fm) + col1[3] * fm
        out.append(oracle.accumulate_linear_const(tseg, c0, c1, al * att))
    return out


class TetOnRay:
    """One tet crossed by a ray: faces, hit points and barycentrics at the original geometry, and its
    sub-segments as functions of the mesh's colours (geometry frozen) or positions (the two frozen-geometry
    derivatives the reference adds, as in Ray above)."""

    def __init__(self, ro, rd, vid, P, C, att):
        self.ro, self.rd, self.vid, self.att = ro, rd, list(vid), att
        Pt = P[self.vid]
        self.hit, self.f0, self.f1, self.t0, self.t1 = oracle.ray_tet(ro, rd, Pt)
        if not self.hit:
            return
        self.faces = [[self.vid[i] for i in FT[self.f0]], [self.vid[i] for i in FT[self.f1]]]
        self.H = [ro + self.t0 * rd, ro + self.t1 * rd]
        self.P0 = P.copy()
        self.C0 = C.copy()
        self.lam = [self._lam(P, self.faces[k], self.H[k]) for k in range(2)]
        self.seg = self.segs_colours(C)

    @staticmethod
    def _lam(P, face, Hk):
        u, v = _bary(P[face[0]], P[face[1]], P[face[2]], Hk)
        return np.array([u, v, 1.0 - u - v])

    def segs_colours(self, C):
        cols = [self.lam[k] @ C[self.faces[k]] for k in range(2)]
        return _subsegs((self.t1 - self.t0) / 4.0, cols[0], cols[1], self.att)

    def segs_positions_i(self, P):
        pts = [self.lam[k] @ P[self.faces[k]] for k in range(2)]
        cols = [self.lam[k] @ self.C0[self.faces[k]] for k in range(2)]
        return _subsegs(np.linalg.norm(pts[0] - pts[1]) / 4.0, cols[0], cols[1], self.att)

    def segs_positions_ii(self, P):
        lam = [self._lam(P, self.faces[k], self.H[k]) for k in range(2)]
        pts = [lam[k] @ self.P0[self.faces[k]] for k in range(2)]
        cols = [lam[k] @ self.C0[self.faces[k]] for k in range(2)]
        return _subsegs(np.linalg.norm(pts[0] - pts[1]) / 4.0, cols[0], cols[1], self.att)

    def stable(self, P):
        """Hit faces unchanged a tiny turn of the ray away, hit points away from the face edges (where the
        unsigned-area barycentrics have their kink)."""
        for d in (np.array([1e-4, 0, 0]), np.array([0, 1e-4, 0]), np.array([0, 0, 1e-4])):
            if oracle.ray_tet(self.ro, self.rd + d, P[self.vid])[1:3] != (self.f0, self.f1):
                return False
        return min(self.lam[0].min(), self.lam[1].min()) >= 0.02


class ChainRay:
    """The frozen parts of the closed form for every tet of one ray."""

    def __init__(self, tets, dl, damp=True, q_with_s0=False):
        self.tets, self.dl = tets, dl
        n = len(tets)
        self.Y, self.Q = [], []
        T = 1.0
        for k, tk in enumerate(tets):
            d = np.prod([1.0 - tets[i].seg[0][3] for i in range(k + 1, n)]) if damp else 1.0
            self.Y.append(np.array([0.0, 0.0, 0.0, 1.0 - T * (1.0 - tk.seg[0][3]) * d]))
            q = np.zeros(4)
            for i in range(k + 1, n):
                for j in range(0 if q_with_s0 else 1, 4):
                    q = _over(q, tets[i].seg[j])
            self.Q.append(q)
            T *= np.prod([1.0 - tk.seg[j][3] for j in range(4)])

    def terms(self, segs_of):
        """(vertex ids, function of the parameters) of each tet's term: the FD of a vertex needs only the terms
        of the tets that hold it."""
        out = []
        for k, tk in enumerate(self.tets):
            def term(x, k=k, tk=tk):
                s = segs_of(tk, x)
                z = _over(_over(_over(self.Y[k], s[1]), s[2]), s[3])
                return float(self.dl @ _over(z, self.Q[k]))
            out.append((set(tk.vid), term))
        return out


def _fd_terms(terms, x0, h):
    """Central differences of the sum of the terms, each differenced only in the vertices it holds."""
    g = np.zeros_like(x0)
    for v in range(x0.shape[0]):
        mine = [fn for vids, fn in terms if v in vids]
        for c in range(x0.shape[1]):
            xp, xm = x0.copy(), x0.copy()
            xp[v, c] += h
            xm[v, c] -= h
            g[v, c] = sum(fn(xp) - fn(xm) for fn in mine) / (2 * h)
    return g


def test_reference_adjoint_chaining_multi_tet_rays():
    w = meshgen.grid_workload(2, seed=5, jitter=0.2)
    pos32 = (w["pos"] * np.float32(0.6)).astype(np.float32)
    cells = w["cells"]
    rng = np.random.default_rng(8)
    col32 = np.concatenate([rng.uniform(0.1, 0.9, size=(pos32.shape[0], 3)),
                            rng.uniform(0.5, 3.0, size=(pos32.shape[0], 1))], axis=1).astype(np.float32)
    W = H = 40
    eye = (0.35, 0.45, 0.75)
    att = 1.0
    uf, _ = oracle.camera_uniforms(view_from(eye), FOVY, ZNEAR, ZFAR, W, H)
    fr = oracle.Fragments(cells, pos32, uf, ZNEAR, W, H)
    img, _, cnt = oracle.forward(fr, pos32, col32, uf, att)
    off, ids = fr.lists()
    P = pos32.astype(np.float64)
    C = col32.astype(np.float64)
    ro = uf[64:67].astype(np.float64)
    dimg = np.zeros((H, W, 4), np.float32)
    rays = []
    nfrag = {2: 0, 3: 0, 4: 0}
    for y in range(H):
        for x in range(W):
            n = int(cnt[y, x])
            if n < 2 or n > 4 or (x + y) % 3:
                continue
            p = y * W + x
            rd = oracle.ray_dir(uf, W, H, x, y).astype(np.float64)
            tets = [TetOnRay(ro, rd, cells[i >> 2], P, C, att) for i in ids[off[p]:off[p + 1]]]
            if not all(t.hit and t.stable(P) for t in tets):
                continue
            dimg[y, x] = rng.normal(size=4)
            rays.append((tets, dimg[y, x].astype(np.float64)))
            nfrag[n] += 1
    print(f"chained rays: {len(rays)} ({nfrag} by fragment count)", flush=True)
    assert len(rays) >= 40 and nfrag[3] + nfrag[4] >= 20
    # the restated sub-segments composite to the oracle's FP64 image on these rays
    img64, _, _ = oracle.forward(fr, P, C, uf.astype(np.float64), att, dtype=np.float64)
    for y in range(H):
        for x in range(W):
            if dimg[y, x].any():
                p = y * W + x
                rd = oracle.ray_dir(uf, W, H, x, y).astype(np.float64)
                acc = np.zeros(4)
                for i in ids[off[p]:off[p + 1]]:
                    for s in TetOnRay(ro, rd, cells[i >> 2], P, C, att).seg:
                        acc = _over(acc, s)
                assert np.abs(acc - img64[y, x]).max() < 1e-6
    gpos, gcol = oracle.adjoint_reference(fr, pos32, col32, uf, att, img, dimg)

    def fd_colours(chains):
        return _fd_terms([t for ch in chains for t in ch.terms(lambda tk, x: tk.segs_colours(x))], C, 1e-6)

    chains = [ChainRay(t, d) for t, d in rays]
    fd_col = fd_colours(chains)
    # positions: the sum of the two frozen-geometry parts' derivatives
    fd_pos = _fd_terms([t for ch in chains for t in ch.terms(lambda tk, x: tk.segs_positions_i(x))] +
                       [t for ch in chains for t in ch.terms(lambda tk, x: tk.segs_positions_ii(x))], P, 1e-7)
    ec, ep = _rel(gcol, fd_col), _rel(gpos, fd_pos)
    print(f"reference adjoint vs the chained closed form: colours rel {ec:.2e}, positions rel {ep:.2e}", flush=True)
    assert ec <= 1e-5
    assert ep <= 1e-4
    # what the closed form rules out: no damping by the sub-segments 0 behind, or every sub-segment inverted
    for kw in (dict(damp=False), dict(damp=True, q_with_s0=True)):
        alt = _rel(gcol, fd_colours([ChainRay(t, d, **kw) for t, d in rays]))
        print(f"  variant {kw}: colours rel {alt:.2e}", flush=True)
        assert alt > 100 * 1e-5
