"""CPU oracle for the learned prescreen (numpy, float64).

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's checks import it; the
product never does). There is no reference counterpart: the reference has no
learned component (Cargo.toml:4). What pins this oracle: finite differences
of its own float64 loss (tests/test_mlp_oracle.py), closed forms of the map
definition (a single antigen atom: the LJ minimum -eps at r = R_ij, Coulomb
332 q / r, the hydrophobic step, a Gaussian peak of 1 at r = mu_k), the exact
all-atom energy that the trilinear map energy converges to at grid points
(tests/test_mlp_oracle.py), and bf16 rounding checked against numpy's
float32 bit patterns.

Maps (maps.hip): at grid point p (float32 origin + h * index) the sums over
antigen atoms i passing the float32 pair predicate 0.25 <= r^2 <= rc^2 (r^2 =
(dx*dx + dy*dy) + dz*dz in float32) of, per LJ class c, sqrt(eps_c eps_i)
(s^12 - 2 s^6) with s = (Rmin/2_c + Rmin/2_i) / r; 332 q_i / r; -0.5 for
hydrophobic i with r^2 < 36; exp(-(r - mu_k)^2 / (2 sigma^2)). Stored as
fp16 after clamping to +-60000; the radial densities are kept on every
second grid point only (an atom reads the nearest of those).

Model: x -> SiLU(x W1 + b1) -> SiLU(. W2 + b2) -> SiLU(. W3 + b3) -> . w4 + b4.
With bf16=True the oracle rounds at the same points as the device path:
weights, the input rows, the hidden activations A1..A3 and the backward
pre-activation gradients dZ1..dZ3 are bf16; everything else is float64.
"""
from __future__ import annotations

import numpy as np

D, H = 32, 256
MAP_H = 0.625
RBF_MU0, RBF_DMU, RBF_SIGMA, N_RBF = 2.0, 1.0, 0.75, 8
MAP_CLAMP = 60000.0
CLASH_E, CONTACT_E = 1.0, -0.05
N_LJ = 7
F_COUL_POS, F_COUL_NEG, F_HYD, F_TOTAL, F_CLASH, F_CONTACT, F_NEAR = 8, 9, 10, 11, 12, 13, 14
F_RBF_POLAR, F_RBF_APOLAR = 16, 24
POLAR_CLASSES = (2, 3, 4, 5)   # N*, O/O2, OH, OS


def bf16_round(a) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    r = np.where(nan, (u >> 16).astype(np.uint32) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(a) -> np.ndarray:
    f = np.ascontiguousarray(bf16_round(a).astype(np.float32))
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def bf16_to_f64(bits) -> np.ndarray:
    return (np.asarray(bits, np.uint32) << 16).view(np.float32).astype(np.float64)


def lj_classes(amber_nb):
    """The distinct (Rmin/2, eps) pairs of the AtomType table, in AtomType order."""
    out = []
    for rmh, eps in amber_nb:
        t = (np.float32(rmh), np.float32(eps))
        if t not in out:
            out.append(t)
    return out


def map_grid(bbox_lo, bbox_hi, cutoff, h=MAP_H):
    """Origin (float32) and points per axis of an antigen's map grid."""
    lo = np.asarray(bbox_lo, np.float32)
    hi = np.asarray(bbox_hi, np.float32)
    pad = np.float32(cutoff) + np.float32(2.0) * np.float32(h)
    origin = (lo - pad).astype(np.float32)
    n = np.ceil(((hi - lo) + np.float32(2.0) * pad) / np.float32(h)).astype(np.int64) + 1
    return origin, n


def map_values(ag, ijk, origin, h, cutoff, classes, hydrophobic_flag, round16=True):
    """Map values [n, 17] at grid points ijk [n, 3] (float64 terms, float32
    predicate, fp16 rounding). ag: dict of float32 arrays x, y, z, charge,
    r_min_half, epsilon and hydrophobic (0/1)."""
    ijk = np.asarray(ijk, np.int64).reshape(-1, 3)
    h32 = np.float32(h)
    P = (origin[None, :] + h32 * ijk.astype(np.float32)).astype(np.float32)
    A = np.stack([ag["x"], ag["y"], ag["z"]], 1).astype(np.float32)
    rc2 = np.float32(cutoff) * np.float32(cutoff)
    q332 = (np.float32(332.0) * ag["charge"].astype(np.float32)).astype(np.float64)
    rmh = ag["r_min_half"].astype(np.float64)
    eps = ag["epsilon"].astype(np.float64)
    hyd = np.asarray(hydrophobic_flag, bool)
    out = np.zeros((len(P), 3 + len(classes) + N_RBF))
    mu = RBF_MU0 + RBF_DMU * np.arange(N_RBF)
    for q in range(len(P)):
        d = P[q][None, :] - A
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        sel = (r2 <= rc2) & (r2 >= np.float32(0.25))
        r2d = r2[sel].astype(np.float64)
        r = np.sqrt(r2d)
        for c, (rc_, ec) in enumerate(classes):
            s2 = (float(rc_) + rmh[sel]) ** 2 / r2d
            s6 = s2 ** 3
            out[q, c] = (np.sqrt(float(ec) * eps[sel]) * (s6 * s6 - 2.0 * s6)).sum()
        out[q, len(classes)] = (q332[sel] / r).sum()
        out[q, len(classes) + 1] = -0.5 * (hyd[sel] & (r2[sel] < np.float32(36.0))).sum()
        out[q, len(classes) + 2:len(classes) + 2 + N_RBF] = np.exp(
            -(r[:, None] - mu[None, :]) ** 2 / (2 * RBF_SIGMA ** 2)).sum(0)
        out[q, -1] = 1.0 if sel.any() else 0.0   # an antigen atom passes the predicate
    if round16:
        out[:, :-1] = fp16_round(out[:, :-1])
    return out


def fp16_round(v):
    return np.clip(np.asarray(v, np.float64), -MAP_CLAMP, MAP_CLAMP).astype(np.float16).astype(np.float64)


def atom_class(rmh, eps, classes):
    """LJ class of an atom: its (Rmin/2, eps) pair, else the nearest one."""
    for c, (r, e) in enumerate(classes):
        if np.float32(rmh) == r and np.float32(eps) == e:
            return c
    d = [abs(float(np.float32(rmh)) - float(r)) + abs(float(np.float32(eps)) - float(e))
         for r, e in classes]
    return int(np.argmin(d))


def atom_map_lookup(x, y, z, origin, h, dims):
    """Cell (i, j, k), fractions (float32 arithmetic as the kernel) and the
    nearest point of the radial-density grid (every second point, given as a
    point of the full grid) of one atom; None outside the map."""
    inv = np.float32(1.0) / np.float32(h)
    u = ((np.array([x, y, z], np.float32) - origin) * inv).astype(np.float32)
    f = np.floor(u)
    if not (np.all(f >= 0) and np.all(f < (np.asarray(dims) - 1))):
        return None
    frac = (u - f).astype(np.float32)
    i = f.astype(np.int64)
    uc = (u * np.float32(0.5)).astype(np.float32)
    fc = np.floor(uc)
    nc = (np.asarray(dims) - 1) // 2 + 1
    near = np.minimum(fc.astype(np.int64) + ((uc - fc) >= np.float32(0.5)), nc - 1)
    return i, frac.astype(np.float64), 2 * near


def trilinear(corners, frac):
    """corners[8] (x fastest) at fractions frac (the kernel's formula)."""
    fx, fy, fz = frac
    gx, gy, gz = 1 - fx, 1 - fy, 1 - fz
    v = corners
    x00 = gx * v[0] + fx * v[1]
    x10 = gx * v[2] + fx * v[3]
    x01 = gx * v[4] + fx * v[5]
    x11 = gx * v[6] + fx * v[7]
    return gz * (gy * x00 + fy * x10) + fz * (gy * x01 + fy * x11)


CORNERS = np.array([[c & 1, (c >> 1) & 1, c >> 2] for c in range(8)])


def features(atoms, origin, h, dims, lookup_values, classes):
    """Features of one candidate (float64 before bf16). atoms: dict of the
    candidate's x, y, z, charge, r_min_half, epsilon, hydrophobic; the map
    grid (origin, spacing, points per axis); lookup_values(ijk [m, 3]) ->
    [m, 18] map values at grid points (the last, at the cell's lower corner,
    the cell's occupancy: an atom in a cold cell adds nothing, nor does one
    outside the map).
    Returns (row[32] transformed, total map energy)."""
    raw = np.zeros(D)
    n = len(atoms["x"])
    for a in range(n):
        lk = atom_map_lookup(atoms["x"][a], atoms["y"][a], atoms["z"][a], origin, h, dims)
        if lk is None:
            continue
        i, frac, near = lk
        if lookup_values(i[None, :])[0, -1] == 0:
            continue                                # a cold cell: zeros only
        lc = atom_class(atoms["r_min_half"][a], atoms["epsilon"][a], classes)
        hydp = atoms["hydrophobic"][a] != 0
        fc = N_LJ if hydp else lc
        vals = lookup_values(i[None, :] + CORNERS)
        elj = trilinear(vals[:, lc], frac)
        phi = trilinear(vals[:, N_LJ], frac)
        hh = trilinear(vals[:, N_LJ + 1], frac)
        rb = lookup_values(near[None, :])[0, N_LJ + 2:N_LJ + 2 + N_RBF]
        q = float(atoms["charge"][a])
        ec = q * phi
        eh = hh if hydp else 0.0
        raw[fc] += elj
        if q > 0:
            raw[F_COUL_POS] += ec
        elif q < 0:
            raw[F_COUL_NEG] += ec
        raw[F_HYD] += eh
        raw[F_TOTAL] += elj + ec + eh
        raw[F_CLASH] += elj > CLASH_E
        raw[F_CONTACT] += elj < CONTACT_E
        raw[F_NEAR] += 1
        if fc in POLAR_CLASSES:
            raw[F_RBF_POLAR:F_RBF_POLAR + N_RBF] += rb
        else:
            raw[F_RBF_APOLAR:F_RBF_APOLAR + N_RBF] += rb
    row = np.zeros(D)
    row[:F_HYD + 1] = np.arcsinh(raw[:F_HYD + 1] * 0.5)
    row[F_TOTAL] = np.arcsinh(raw[F_TOTAL] * 0.1)
    row[F_CLASH:F_NEAR + 1] = np.log1p(raw[F_CLASH:F_NEAR + 1])
    row[F_RBF_POLAR:] = np.log1p(raw[F_RBF_POLAR:])
    return row, raw[F_TOTAL]


def unpack(p, d=D, h=H):
    o = 0
    out = {}
    for name, shape in (("W1", (d, h)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)),
                        ("W3", (h, h)), ("b3", (h,)), ("w4", (h,)), ("b4", (1,))):
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    assert o == len(p)
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def loss_grad(p, x, t, d=D, h=H, bf16=True):
    """Mean squared error and its gradient w.r.t. the flat parameters."""
    P = unpack(np.asarray(p, np.float64), d, h)
    r = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    Wq = {k: (r(v.astype(np.float32)) if bf16 and k in ("W1", "W2", "W3", "w4") else v)
          for k, v in P.items()}
    X = np.asarray(x, np.float64)
    B = X.shape[0]
    Z1 = X @ Wq["W1"] + P["b1"]
    A1 = r(silu(Z1).astype(np.float32)) if bf16 else silu(Z1)
    Z2 = A1 @ Wq["W2"] + P["b2"]
    A2 = r(silu(Z2).astype(np.float32)) if bf16 else silu(Z2)
    Z3 = A2 @ Wq["W3"] + P["b3"]
    A3 = r(silu(Z3).astype(np.float32)) if bf16 else silu(Z3)
    y = A3 @ Wq["w4"] + P["b4"][0]
    res = y - np.asarray(t, np.float64)
    loss = float((res ** 2).mean())
    dy = 2.0 * res / B
    g = {"w4": A3.T @ dy, "b4": np.array([dy.sum()])}
    q = (lambda v: r(v.astype(np.float32))) if bf16 else (lambda v: v)
    dZ3 = q(np.outer(dy, Wq["w4"]) * dsilu(Z3))
    g["W3"] = A2.T @ dZ3
    g["b3"] = dZ3.sum(0)
    dZ2 = q((dZ3 @ Wq["W3"].T) * dsilu(Z2))
    g["W2"] = A1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = q((dZ2 @ Wq["W2"].T) * dsilu(Z1))
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    flat = np.concatenate([g[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4",
                                                       "b4")])
    return loss, flat, y


def predict(p, x, d=D, h=H, bf16=True):
    return loss_grad(p, x, np.zeros(len(x)), d, h, bf16)[2]
