"""Coil geometry helpers: recognise a coil region's shape from the mesh and
build its unit current-direction field (src/utils/CoilGeometryUtils.m:1-581).

Three shapes, each a detector and a direction field:
- circular (:17-80): centre, skeleton radius, cross-section area, axis;
- racetrack (:82-166): two straight sides joined by half circles;
- rounded rectangle (:169-271): four straight sides, four quarter circles.

All detectors share ``detectGeneralRect`` (:273-372), which measures a coil
lying in a plane:
1. The coil's normal comes from PCA of its nodes (``fit_circle_normal_robust``,
   :374-421): the eigenvector whose eigenvalue is set apart from the other two.
2. The in-plane axes come from the minimum-area bounding rectangle
   (a bounded scalar minimisation over [0, pi/2]), U the longer side.
3. The height H is the spread along the normal.
4. Two cutting planes through the centre, normal to V and to U, are intersected
   with the coil's tets (``computeSliceProperties``, :423-516). The cut area
   gives the wall width (area / 2H). The distance between the area centroids
   on either side gives the skeleton span along the measured axis.
5. The corner radius is the skeleton radius that best fits the corner nodes
   (RMS distance to the skeleton outline).

Host-side numpy: this runs once per coil, before any assembly. Arrays follow
this package's mesh layout: ``mesh.P`` (Nn, 3), ``mesh.T`` (Ne, 4) 1-based.
Points, centres and directions are returned in the reference's layout:
3-vectors, and (3, Ne) for direction fields.
``fminbnd`` is MATLAB's bounded Brent minimiser (golden section with
parabolic steps). scipy's ``minimize_scalar(method="bounded")`` implements
the same method, with the same absolute tolerance on x.
"""
from __future__ import annotations

import numpy as np

# tet edges in the order the slice routine visits them (CoilGeometryUtils.m:456)
_TET_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def _fminbnd(fun, lo: float, hi: float, tol_x: float = 1e-4) -> float:
    from scipy.optimize import minimize_scalar
    res = minimize_scalar(fun, bounds=(lo, hi), method="bounded", options={"xatol": tol_x})
    return float(res.x)


class CoilGeometryUtils:
    # ------------------------------------------------------------- helpers
    @staticmethod
    def getPts(mesh, regionID):
        """Unique nodes of a region (sorted ids) and their coordinates (3, Np)
        (CoilGeometryUtils.m:541-547)."""
        mask = np.asarray(mesh.RegionTags) == regionID
        if not mask.any():
            raise ValueError(f"Region {regionID} is empty!")
        nodes = np.unique(np.asarray(mesh.T)[mask].ravel())
        return np.asarray(mesh.P)[nodes - 1].T.copy(), nodes

    @staticmethod
    def getElementCenters(mesh) -> np.ndarray:
        """(3, Ne) tet centroids (:554-557)."""
        P = np.asarray(mesh.P)
        T = np.asarray(mesh.T) - 1
        return ((P[T[:, 0]] + P[T[:, 1]] + P[T[:, 2]] + P[T[:, 3]]) / 4.0).T

    @staticmethod
    def getUVIndices(axis_idx: int):
        """0-based in-plane axes for a 1-based normal axis (:559-564)."""
        if axis_idx == 1:
            return 2, 1
        if axis_idx == 2:
            return 0, 2
        return 0, 1

    @staticmethod
    def bboxArea(theta, u0, v0) -> float:
        c, s = np.cos(theta), np.sin(theta)
        ur = c * u0 + s * v0
        vr = -s * u0 + c * v0
        return float((ur.max() - ur.min()) * (vr.max() - vr.min()))

    @staticmethod
    def skeletonFitError(R, u, v, X_lim, Y_lim) -> float:
        """RMS distance of corner points to a rounded-rectangle skeleton of
        half-spans X_lim, Y_lim and corner radius R (:525-539)."""
        cx, cy = X_lim - R, Y_lim - R
        dist = np.zeros_like(u)
        top = u <= cx
        dist[top] = np.abs(v[top] - Y_lim)
        right = (v <= cy) & ~top
        dist[right] = np.abs(u[right] - X_lim)
        arc = ~top & ~right
        if arc.any():
            dx, dy = u[arc] - cx, v[arc] - cy
            dist[arc] = np.abs(np.sqrt(dx * dx + dy * dy) - R)
        return float(np.sqrt(np.mean(dist * dist)))

    @staticmethod
    def fit_circle_normal_robust(pts):
        """Normal of a planar coil from PCA of its points (3, N) (:374-421):
        the eigenvector whose eigenvalue is farthest from the middle one, its
        largest component made positive.  Returns (normal, centre)."""
        center = pts.mean(axis=1)
        pc = pts - center[:, None]
        cov = pc @ pc.T / (pc.shape[1] - 1)
        d, V = np.linalg.eigh(cov)  # ascending
        axis = 2 if abs(d[1] - d[0]) < abs(d[2] - d[1]) else 0
        normal = V[:, axis].copy()
        k = int(np.argmax(np.abs(normal)))
        if normal[k] < 0:
            normal = -normal
        return normal, center

    @staticmethod
    def computeSliceProperties(mesh, regionID, center, plane_n, measure_axis):
        """Area of the region's cut by the plane through `center` with normal
        `plane_n`, and the area centroids' coordinates along `measure_axis`
        on its positive and negative sides (:423-516); None for an empty side.

        Each tet that straddles the plane contributes the polygon whose
        corners are the edge crossings, taken in edge order (an edge crosses
        when its end signs differ, a zero sign included).  The corners are
        sorted by angle about their mean and fanned from it.  A fourth corner
        within 1e-6 of the first is dropped."""
        P = np.asarray(mesh.P)
        mask = np.asarray(mesh.RegionTags) == regionID
        Tsub = np.asarray(mesh.T)[mask] - 1
        D_all = (P - center[None, :]) @ plane_n
        De = D_all[Tsub]
        cross_m = (De.min(axis=1) < -1e-9) & (De.max(axis=1) > 1e-9)
        Tc = Tsub[cross_m]
        if Tc.shape[0] == 0:
            return 0.0, None, None
        d = D_all[Tc]                      # (Nc, 4)
        p = P[Tc]                          # (Nc, 4, 3)
        ia = np.array([e[0] for e in _TET_EDGES])
        ib = np.array([e[1] for e in _TET_EDGES])
        d1, d2 = d[:, ia], d[:, ib]        # (Nc, 6)
        hit = np.sign(d1) != np.sign(d2)
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.where(hit, -d1 / np.where(hit, d2 - d1, 1.0), 0.0)
        pts = p[:, ia] + t[..., None] * (p[:, ib] - p[:, ia])   # (Nc, 6, 3)
        cnt = hit.sum(axis=1)
        # crossings packed to the front in edge order
        order = np.argsort(~hit, axis=1, kind="stable")
        pts = np.take_along_axis(pts, order[..., None], axis=1)
        dup = (cnt == 4) & (np.linalg.norm(pts[:, 0] - pts[:, 3], axis=1) < 1e-6)
        cnt = np.where(dup, 3, cnt)
        v1 = measure_axis
        v2 = np.cross(plane_n, measure_axis)
        area_all = np.zeros(Tc.shape[0])
        cpoly_all = np.zeros((Tc.shape[0], 3))
        for c in range(3, 7):
            sel = np.nonzero(cnt == c)[0]
            if sel.size == 0:
                continue
            q = pts[sel, :c]                                   # (m, c, 3)
            cpoly = q.sum(axis=1) / c
            vec = q - cpoly[:, None, :]
            ang = np.arctan2(vec @ v2, vec @ v1)
            srt = np.argsort(ang, axis=1, kind="stable")
            q = np.take_along_axis(q, srt[..., None], axis=1)
            a = q - cpoly[:, None, :]
            b = np.roll(q, -1, axis=1) - cpoly[:, None, :]
            area = 0.5 * np.linalg.norm(np.cross(a, b), axis=2)   # (m, c)
            # the reference adds the fan triangles one after another
            tot = np.zeros(sel.size)
            for k in range(c):
                tot = tot + area[:, k]
            area_all[sel] = tot
            cpoly_all[sel] = cpoly
        keep = cnt >= 3
        proj = (cpoly_all - center[None, :]) @ measure_axis
        total = 0.0
        w_pos = a_pos = w_neg = a_neg = 0.0
        # running sums in tet order, as the reference's loop accumulates them
        for i in np.nonzero(keep)[0]:
            ar = area_all[i]
            total += ar
            if proj[i] > 0:
                w_pos += proj[i] * ar
                a_pos += ar
            else:
                w_neg += proj[i] * ar
                a_neg += ar
        cpos = w_pos / a_pos if a_pos > 0 else None
        cneg = w_neg / a_neg if a_neg > 0 else None
        return total, cpos, cneg

    @staticmethod
    def detectGeneralRect(mesh, regionID):
        """Returns (center, H, Lx_mean, Ly_mean, R_mean, normal_axis_idx,
        WallWidths, axes) of a planar coil (:273-372); normal_axis_idx is
        1-based like the reference's; ``axes`` = (axis_u, axis_v, axis_n)."""
        pts, _ = CoilGeometryUtils.getPts(mesh, regionID)
        axis_n, center = CoilGeometryUtils.fit_circle_normal_robust(pts)
        pc = pts - center[:, None]
        normal_axis_idx = int(np.argmax(np.abs(axis_n))) + 1
        base_v = np.cross(axis_n, [1.0, 0.0, 0.0])
        if np.linalg.norm(base_v) < 1e-3:
            base_v = np.cross(axis_n, [0.0, 1.0, 0.0])
        base_v = base_v / np.linalg.norm(base_v)
        base_u = np.cross(axis_n, base_v)
        u0, v0 = base_u @ pc, base_v @ pc
        theta = _fminbnd(lambda th: CoilGeometryUtils.bboxArea(th, u0, v0), 0.0, np.pi / 2)
        c, s = np.cos(theta), np.sin(theta)
        au = c * base_u + s * base_v
        av = -s * base_u + c * base_v
        ut, vt = np.abs(au @ pc), np.abs(av @ pc)
        if (vt.max() - vt.min()) > (ut.max() - ut.min()):
            axis_u, axis_v = av, au
        else:
            axis_u, axis_v = au, av
        w = axis_n @ pc
        H = float(w.max() - w.min())
        if H < 1e-9:
            H = 1.0
        area_u, cu_pos, cu_neg = CoilGeometryUtils.computeSliceProperties(
            mesh, regionID, center, axis_v, axis_u)
        ww_u = area_u / (2 * H)
        if cu_pos is None or cu_neg is None:
            Lx_mean = float(np.abs(axis_u @ pc).max() * 2 - ww_u)
        else:
            Lx_mean = abs(cu_pos - cu_neg)
        area_v, cv_pos, cv_neg = CoilGeometryUtils.computeSliceProperties(
            mesh, regionID, center, axis_u, axis_v)
        ww_v = area_v / (2 * H)
        if cv_pos is None or cv_neg is None:
            Ly_mean = float(np.abs(axis_v @ pc).max() * 2 - ww_v)
        else:
            Ly_mean = abs(cv_pos - cv_neg)
        u, v = np.abs(axis_u @ pc), np.abs(axis_v @ pc)
        X_lim, Y_lim = Lx_mean / 2, Ly_mean / 2
        corner = (u > X_lim * 0.5) & (v > Y_lim * 0.5)
        if corner.sum() < 10:
            R_mean = 0.0
        else:
            uc, vc = u[corner], v[corner]
            R_mean = _fminbnd(lambda r: CoilGeometryUtils.skeletonFitError(r, uc, vc, X_lim, Y_lim),
                              0.0, min(X_lim, Y_lim))
            if R_mean < 1e-3:
                R_mean = 0.0
        return (center, H, Lx_mean, Ly_mean, R_mean, normal_axis_idx,
                np.array([ww_u, ww_v]), (axis_u, axis_v, axis_n))

    # ------------------------------------------------------------- circular
    @staticmethod
    def autoDetectCircular(mesh, regionID):
        """(center, R, area, normal_axis_idx) (:17-35)."""
        center, H, lx, ly, _, ax, ww, _ = CoilGeometryUtils.detectGeneralRect(mesh, regionID)
        R = (lx + ly) / 2 / 2
        return center, R, float(ww.mean() * H), ax

    @staticmethod
    def computeCircularDirection(mesh, regionID, center, axis_idx) -> np.ndarray:
        """(3, Ne) unit tangents (-v, u)/|.| about `center` in the coil's
        plane; zero outside the region and on the axis (:37-80)."""
        cen = CoilGeometryUtils.getElementCenters(mesh)
        mask = np.asarray(mesh.RegionTags) == regionID
        out = np.zeros((3, mesh.NumElements))
        if not mask.any():
            return out
        loc = cen[:, mask] - np.asarray(center, dtype=np.float64).reshape(3, 1)
        ui, vi = CoilGeometryUtils.getUVIndices(axis_idx)
        u, v = loc[ui], loc[vi]
        mag = np.sqrt(u * u + v * v)
        ok = mag > 1e-12
        idx = np.nonzero(mask)[0][ok]
        out[ui, idx] = -v[ok] / mag[ok]
        out[vi, idx] = u[ok] / mag[ok]
        return out

    # ------------------------------------------------------------ racetrack
    @staticmethod
    def autoDetectRacetrack(mesh, regionID):
        """(center, L_straight, R_curve, area, normal_axis_idx) (:82-108)."""
        center, H, lx, ly, rm, ax, ww, _ = CoilGeometryUtils.detectGeneralRect(mesh, regionID)
        L = (lx if lx > ly else ly) - 2 * rm
        return center, max(L, 0.0), rm, float(ww.mean() * H), ax

    @staticmethod
    def computeRacetrackDirection(mesh, regionID, center, straight_len, axis_idx) -> np.ndarray:
        """(3, Ne): along -sign(v) U on the straight part |u| <= L/2, circular
        tangents about (+-L/2, 0) on the ends (:110-166)."""
        cen = CoilGeometryUtils.getElementCenters(mesh)
        mask = np.asarray(mesh.RegionTags) == regionID
        out = np.zeros((3, mesh.NumElements))
        if not mask.any():
            return out
        full = np.nonzero(mask)[0]
        loc = cen[:, mask] - np.asarray(center, dtype=np.float64).reshape(3, 1)
        ui, vi = CoilGeometryUtils.getUVIndices(axis_idx)
        u, v = loc[ui], loc[vi]
        st = np.abs(u) <= straight_len / 2
        out[ui, full[st]] = -np.sign(v[st])
        out[vi, full[st]] = 0.0
        cu = ~st
        if cu.any():
            us = u[cu] - np.sign(u[cu]) * (straight_len / 2)
            vc = v[cu]
            mag = np.sqrt(us * us + vc * vc)
            ok = mag > 1e-12
            idx = full[np.nonzero(cu)[0][ok]]
            out[ui, idx] = -vc[ok] / mag[ok]
            out[vi, idx] = us[ok] / mag[ok]
        return out

    # ----------------------------------------------------- rounded rectangle
    @staticmethod
    def autoDetectRoundedRect(mesh, regionID):
        """(center, R_in, R_out, Lx, Ly, area, normal_axis_idx) (:169-199):
        Lx, Ly the skeleton's straight lengths, area = mean wall width x H."""
        center, H, lxm, lym, rm, ax, ww, _ = CoilGeometryUtils.detectGeneralRect(mesh, regionID)
        Lx, Ly = lxm - 2 * rm, lym - 2 * rm
        wall = float(ww.mean())
        R_out, R_in = rm + wall / 2, rm - wall / 2
        return (center, max(R_in, 0.0), R_out, max(Lx, 0.0), max(Ly, 0.0), wall * H, ax)

    @staticmethod
    def computeRoundedRectDirection(mesh, regionID, center, Lx, Ly, axis_idx) -> np.ndarray:
        """(3, Ne) (:201-271): -sign(v) U on the sides |u| <= Lx/2, sign(u) V
        on the sides |v| <= Ly/2, circular tangents about the corner centres
        (sign(u) Lx/2, sign(v) Ly/2) elsewhere."""
        cen = CoilGeometryUtils.getElementCenters(mesh)
        mask = np.asarray(mesh.RegionTags) == regionID
        out = np.zeros((3, mesh.NumElements))
        if not mask.any():
            return out
        full = np.nonzero(mask)[0]
        loc = cen[:, mask] - np.asarray(center, dtype=np.float64).reshape(3, 1)
        ui, vi = CoilGeometryUtils.getUVIndices(axis_idx)
        u, v = loc[ui], loc[vi]
        tb = np.abs(u) <= Lx / 2
        lr = (np.abs(v) <= Ly / 2) & ~tb
        co = ~tb & ~lr
        out[ui, full[tb]] = -np.sign(v[tb])
        out[vi, full[tb]] = 0.0
        out[ui, full[lr]] = 0.0
        out[vi, full[lr]] = np.sign(u[lr])
        if co.any():
            uc, vc = u[co], v[co]
            du = uc - np.sign(uc) * (Lx / 2)
            dv = vc - np.sign(vc) * (Ly / 2)
            mag = np.sqrt(du * du + dv * dv)
            ok = mag > 1e-12
            idx = full[np.nonzero(co)[0][ok]]
            out[ui, idx] = -dv[ok] / mag[ok]
            out[vi, idx] = du[ok] / mag[ok]
        return out
