"""ORACLE -- test infrastructure only.

CPU restatement (numpy, float32 in the device's operation order) of the warm
start's host-independent rules in bullet3_amd/csrc/k_warm.hip: matching a
constraint to the previous step's constraint on the same body pair, the
initial impulses of the previous-frame mode, the 24 feature rows and the
logged targets.  The reference has no warm start (its solver zeroes the
impulses, b3ConvertConstraint4.h:73,107,132); what pins this file is the
independent restatement plus the invariants in tests/test_gpu_warm.py (a zero
network reproduces the previous-frame mode, warm and cold solves share their
fixed point).

Constraints are numpy arrays of bullet3_amd.layout.CONSTRAINT, bodies of BODY.
"""
import numpy as np

f32 = np.float32
MATCH_TOL = f32(0.1)


def dot3(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def cross3(a, b):
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], f32)


def plane_space(n):
    """b3PlaneSpace1 (two tangents of the unit normal n), float32"""
    n = np.asarray(n[:3], f32)
    if abs(n[2]) > f32(0.70710678):
        a = f32(f32(n[1] * n[1]) + f32(n[2] * n[2]))
        k = f32(1.0) / np.sqrt(a)
        p = np.array([0.0, -n[2] * k, n[1] * k], f32)
        q = np.array([a * k, -n[0] * p[2], n[0] * p[1]], f32)
    else:
        a = f32(f32(n[0] * n[0]) + f32(n[1] * n[1]))
        k = f32(1.0) / np.sqrt(a)
        p = np.array([-n[1] * k, n[0] * k, 0.0], f32)
        q = np.array([-n[2] * p[1], n[2] * p[0], a * k], f32)
    return p, q


def num_points(c):
    return int(np.count_nonzero(c["jacCoeffInv"] != 0))


def match(prev_cons, cons):
    """per current constraint: matched previous impulses [n][8] =
    (normal 4, friction 2 on the current tangents, pair matched, points)"""
    index = {}
    for j, c in enumerate(prev_cons):
        index[(int(c["bodyA"]), int(c["bodyB"]))] = j
    out = np.zeros((len(cons), 8), f32)
    tol2 = f32(MATCH_TOL * MATCH_TOL)
    for k, c in enumerate(cons):
        npt = num_points(c)
        out[k, 7] = npt
        j = index.get((int(c["bodyA"]), int(c["bodyB"])))
        if j is None:
            continue
        h = prev_cons[j]
        pn = num_points(h)
        for i in range(npt):
            p = c["worldPos"][i, :3]
            best, best_d = -1, tol2
            for m in range(pn):
                d = (p - h["worldPos"][m, :3]).astype(f32)
                dd = dot3(d, d)
                if dd < best_d:
                    best_d, best = dd, m
            if best >= 0:
                out[k, i] = h["appliedRambdaDt"][best]
        t1p, t2p = plane_space(h["linear"])
        fr = (t1p * h["fAppliedRambdaDt"][0] + t2p * h["fAppliedRambdaDt"][1]).astype(f32)
        t1, t2 = plane_space(c["linear"])
        out[k, 4] = dot3(fr, t1)
        out[k, 5] = dot3(fr, t2)
        out[k, 6] = 1.0
    return out


def initial_impulses_previous_frame(prev):
    """the previous-frame mode's initial impulses [n][6] from match()"""
    return np.ascontiguousarray(prev[:, :6], f32)


def point_matched(prev_cons, cons):
    """[n][4] 1 where the point found a previous point"""
    index = {(int(c["bodyA"]), int(c["bodyB"])): j for j, c in enumerate(prev_cons)}
    out = np.zeros((len(cons), 4), f32)
    tol2 = f32(MATCH_TOL * MATCH_TOL)
    for k, c in enumerate(cons):
        j = index.get((int(c["bodyA"]), int(c["bodyB"])))
        if j is None:
            continue
        h = prev_cons[j]
        for i in range(num_points(c)):
            p = c["worldPos"][i, :3]
            for m in range(num_points(h)):
                d = (p - h["worldPos"][m, :3]).astype(f32)
                if dot3(d, d) < tol2:
                    out[k, i] = 1.0
                    break
    return out


def features(cons, bodies, prev, pmatch, vel_unit):
    """feature rows [4n][24] (rows past a constraint's points are zero)"""
    u = f32(vel_unit)
    X = np.zeros((4 * len(cons), 24), f32)
    for k, c in enumerate(cons):
        npt = int(prev[k, 7])
        A, B = bodies[int(c["bodyA"])], bodies[int(c["bodyB"])]
        n = c["linear"][:3].astype(f32)
        t1, t2 = plane_space(c["linear"])
        ims = f32(A["invMass"] + B["invMass"])
        lp = prev[k, :4]
        fnp = f32(max(npt, 1))
        sf = np.abs(c["fJacCoeffInv"]).astype(f32) * u
        dw = (A["angVel"][:3] - B["angVel"][:3]).astype(f32)
        for i in range(npt):
            x = X[4 * k + i]
            p = c["worldPos"][i, :3].astype(f32)
            rA = (p - A["pos"][:3]).astype(f32)
            rB = (p - B["pos"][:3]).astype(f32)
            vA = (A["linVel"][:3] + cross3(A["angVel"], rA)).astype(f32)
            vB = (B["linVel"][:3] + cross3(B["angVel"], rB)).astype(f32)
            dv = (vA - vB).astype(f32)
            si = f32(abs(c["jacCoeffInv"][i]) * u)
            off = (p - c["center"][:3]).astype(f32)
            x[0:3] = n
            x[3] = dot3(dv, n) / u
            x[4] = dot3(dv, t1) / u
            x[5] = dot3(dv, t2) / u
            x[6] = c["b"][i] / u
            x[7] = c["worldPos"][i, 3] / f32(0.05)
            x[8] = A["invMass"] / ims
            x[9] = B["invMass"] / ims
            x[10] = abs(c["jacCoeffInv"][i]) * ims
            x[11] = abs(c["fJacCoeffInv"][0]) * ims
            x[12] = abs(c["fJacCoeffInv"][1]) * ims
            x[13] = lp[i] / si
            x[14] = prev[k, 4] / (fnp * sf[0])
            x[15] = prev[k, 5] / (fnp * sf[1])
            x[16] = prev[k, 6]
            x[17] = pmatch[k, i]
            x[18] = fnp * f32(0.25)
            x[19] = f32(f32(lp[0] + lp[1]) + f32(lp[2] + lp[3])) / (fnp * si)
            x[20] = dot3(off, t1) / f32(0.5)
            x[21] = dot3(off, t2) / f32(0.5)
            x[22] = dot3(dw, n) * f32(0.5) / u
            x[23] = c["linear"][3]
    return X


def targets(cons_after, prev, vel_unit):
    """logged targets [4n][3] from the solve's final impulses: the manifold's
    total normal impulse change shared by its points, the friction changes"""
    u = f32(vel_unit)
    T = np.zeros((4 * len(cons_after), 3), f32)
    for k, c in enumerate(cons_after):
        npt = int(prev[k, 7])
        fnp = f32(max(npt, 1))
        sf = np.abs(c["fJacCoeffInv"]).astype(f32) * u
        t1 = (c["fAppliedRambdaDt"][0] - prev[k, 4]) / (fnp * sf[0])
        t2 = (c["fAppliedRambdaDt"][1] - prev[k, 5]) / (fnp * sf[1])
        tot, tot_prev = f32(0), f32(0)
        for i in range(npt):
            tot = f32(tot + c["appliedRambdaDt"][i])
            tot_prev = f32(tot_prev + prev[k, i])
        for i in range(npt):
            si = f32(abs(c["jacCoeffInv"][i]) * u)
            T[4 * k + i] = ((tot - tot_prev) / (fnp * si), t1, t2)
    return T
