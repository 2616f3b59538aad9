"""Fréchet rows and the misfit gradient of the surface-wave travel times on
the GPU: the row loops of SurfTime::frechet_matrix and SurfTime::compute_grad
(src/SWD/SurfaceWave.cpp:554-676, 695-781) over the 1-D depth kernels that
jdsg_surftime_kernel leaves in HBM.

A data row is one source-receiver ray; its sensitivity fdm(ix, iy) to the
2-D velocity map (the eikonal ray tracer's output) comes in sparse form:
increasing grid cells n = ix + iy*nx and float weights.  Row r is
float(fdm(ix, iy) * kernel(n, iz, idx_r)) over the interior model cells
c = iz*(nx-2)*(ny-2) + (iy-1)*(nx-2) + (ix-1), nonzero entries in increasing
c; the gradient is sum_r row_r * (t_r - t_obs,r), times vs at the cell.
"""
import ctypes as C
import json
import os
import struct
import time

import numpy as np

from . import swd


def _L():
    L = swd.lib()
    vp = C.c_void_p
    L.jdsg_frechet_rows.argtypes = [C.c_int] * 4 + [vp, C.c_int] + [vp] * 4 + \
        [vp, vp, vp, C.c_longlong, C.POINTER(C.c_longlong), vp]
    L.jdsg_frechet_grad.argtypes = [C.c_int] * 4 + [vp, C.c_int] + [vp] * 4 + [vp, vp, vp, vp]
    L.jdsg_write_frechet.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
    return L


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def rays_from_dense(fdm):
    """Sparse form (ray_off, ray_cell, ray_w) of dense sensitivities fdm
    (nrows, nx, ny) with fdm[r, ix, iy]: the nonzero cells n = ix + iy*nx of
    each row in increasing order."""
    nrows, nx, ny = fdm.shape
    flat = fdm.transpose(0, 2, 1).reshape(nrows, nx * ny)   # n = ix + iy*nx
    rr, nn = np.nonzero(flat)
    off = np.zeros(nrows + 1, np.int64)
    np.add.at(off, rr + 1, 1)
    return np.cumsum(off), nn.astype(np.int32), flat[rr, nn].astype(np.float32)


def dense_from_rays(nx, ny, ray_off, ray_cell, ray_w):
    nrows = len(ray_off) - 1
    out = np.zeros((nrows, ny, nx), np.float32)
    for r in range(nrows):
        a, b = ray_off[r], ray_off[r + 1]
        out[r].reshape(-1)[ray_cell[a:b]] = ray_w[a:b]
    return out.transpose(0, 2, 1).copy()


def straight_rays(nx, ny, nrows, seed=0, step=0.25):
    """Synthetic ray sensitivities: straight source-receiver segments across
    the grid (unit node spacing), sampled every `step` and spread bilinearly
    onto the four nodes around each sample (the sparsity of an interpolated
    ray-path Fréchet row).  Returns the sparse form."""
    rng = np.random.default_rng(seed)
    p0 = rng.uniform([0, 0], [nx - 1, ny - 1], (nrows, 2))
    p1 = rng.uniform([0, 0], [nx - 1, ny - 1], (nrows, 2))
    length = np.linalg.norm(p1 - p0, axis=1)
    ns = np.maximum(2, np.ceil(length / step).astype(np.int64))
    offs = np.concatenate([[0], np.cumsum(ns)])
    row = np.repeat(np.arange(nrows), ns)
    t = (np.arange(offs[-1]) - offs[row]) / (ns[row] - 1)
    pts = p0[row] + (p1 - p0)[row] * t[:, None]
    wseg = (length / ns)[row]
    x0 = np.clip(np.floor(pts[:, 0]).astype(np.int64), 0, nx - 2)
    y0 = np.clip(np.floor(pts[:, 1]).astype(np.int64), 0, ny - 2)
    fx, fy = pts[:, 0] - x0, pts[:, 1] - y0
    keys, ws = [], []
    for dx, dy, wt in ((0, 0, (1 - fx) * (1 - fy)), (1, 0, fx * (1 - fy)),
                       (0, 1, (1 - fx) * fy), (1, 1, fx * fy)):
        keys.append(row * (nx * ny) + (y0 + dy) * nx + (x0 + dx))
        ws.append(wt * wseg)
    keys = np.concatenate(keys)
    ws = np.concatenate(ws)
    uk, inv = np.unique(keys, return_inverse=True)
    w = np.zeros(len(uk))
    np.add.at(w, inv, ws)
    keep = w > 0
    uk, w = uk[keep], w[keep]
    r = uk // (nx * ny)
    off = np.zeros(nrows + 1, np.int64)
    np.add.at(off, r + 1, 1)
    return np.cumsum(off), (uk % (nx * ny)).astype(np.int32), (-w / 9.0).astype(np.float32)


def frechet_rows(kernel, nx, ny, nz, kmax, row_idx, ray_off, ray_cell, ray_w, cap=None,
                 stream=None):
    """CSR rows on the GPU.  kernel: CUDA float64 tensor of nx*ny*nz*kmax in
    the (nx*ny, nz, kmax) column-major order; the rest host or CUDA arrays.
    Returns CUDA tensors (row_off int64, col int32, val float32)."""
    import torch
    L = _L()
    dev = [torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a, device="cuda")
           for a in (row_idx, ray_off, ray_cell, ray_w)]
    ri, ro_in, rc, rw = dev[0].to(torch.int32), dev[1].to(torch.int64), dev[2].to(torch.int32), \
        dev[3].to(torch.float32)
    nrows = ri.numel()
    if cap is None:
        cap = int(ro_in[-1].item()) * (nz - 1) if nrows else 0
    row_off = torch.empty(nrows + 1, dtype=torch.int64, device="cuda")
    col = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
    val = torch.empty(max(cap, 1), dtype=torch.float32, device="cuda")
    nnz = C.c_longlong(0)
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    rc_ = L.jdsg_frechet_rows(nx, ny, nz, kmax, _p(kernel), nrows, _p(ri), _p(ro_in), _p(rc),
                              _p(rw), _p(row_off), _p(col), _p(val), cap, C.byref(nnz),
                              C.c_void_p(st))
    if rc_:
        raise RuntimeError(f"jdsg_frechet_rows: {swd.last_error()}")
    return row_off, col[:nnz.value], val[:nnz.value]


def frechet_grad(kernel, nx, ny, nz, kmax, row_idx, ray_off, ray_cell, ray_w, resid, vs,
                 stream=None):
    """Misfit gradient ((nx-2)*(ny-2)*(nz-1) float32 CUDA tensor); vs is the
    (nx, ny, nz) model (host array or CUDA tensor in column-major order)."""
    import torch
    L = _L()

    def dev(a, dt):
        return (a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a))).to("cuda", dt)
    ri, ro, rc, rw = dev(row_idx, torch.int32), dev(ray_off, torch.int64), \
        dev(ray_cell, torch.int32), dev(ray_w, torch.float32)
    res = dev(resid, torch.float32)
    if torch.is_tensor(vs):
        vsd = vs.to("cuda", torch.float32).contiguous()
    else:
        vsd = torch.from_numpy(np.ascontiguousarray(np.asarray(vs, np.float32).ravel(order="F")))
        vsd = vsd.cuda()
    grad = torch.empty((nx - 2) * (ny - 2) * (nz - 1), dtype=torch.float32, device="cuda")
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    rc_ = L.jdsg_frechet_grad(nx, ny, nz, kmax, _p(kernel), ri.numel(), _p(ri), _p(ro), _p(rc),
                              _p(rw), _p(res), _p(vsd), _p(grad), C.c_void_p(st))
    if rc_:
        raise RuntimeError(f"jdsg_frechet_grad: {swd.last_error()}")
    return grad


def write_frechet(path, m, n, row_id, row_off, col, val):
    """frechet.bin as the reference's merge_csr_files leaves it (host arrays)."""
    L = _L()
    a = [np.ascontiguousarray(x, dt) for x, dt in
         ((row_id, np.int32), (row_off, np.int64), (col, np.int32), (val, np.float32))]
    if L.jdsg_write_frechet(os.fsencode(path), m, n, len(a[0]),
                            *[C.c_void_p(x.ctypes.data) for x in a]):
        raise RuntimeError(f"jdsg_write_frechet: {swd.last_error()}")


def read_frechet(path):
    """(m, n, [(row id, cols, vals), ...]) from a frechet.bin file."""
    data = open(path, "rb").read()
    m, n = struct.unpack_from("<ii", data, 0)
    pos, rows = 8, []
    while pos < len(data):
        rid, nar = struct.unpack_from("<ii", data, pos)
        pos += 8
        cols = np.frombuffer(data, np.int32, nar, pos)
        pos += 4 * nar
        vals = np.frombuffer(data, np.float32, nar, pos)
        pos += 4 * nar
        rows.append((rid, cols, vals))
    return m, n, rows


def bench(nx=256, ny=256, nz=17, kmax=120, nrows=20000, steps=5, seed=0):
    """Rows and gradient for synthetic straight rays over a grid256-sized grid
    (256 x 256 columns, 17 depth nodes = 65 layer entries, kmax period
    indices) with random kernels resident in HBM; returns a report dict."""
    import torch
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    kernel = torch.randn(nx * ny * nz * kmax, dtype=torch.float64, device="cuda", generator=g)
    ray_off, ray_cell, ray_w = straight_rays(nx, ny, nrows, seed)
    row_idx = rng.integers(0, kmax, nrows).astype(np.int32)
    dev = [torch.as_tensor(a).cuda() for a in (row_idx, ray_off, ray_cell, ray_w)]
    resid = torch.randn(nrows, device="cuda", generator=g)
    vs = torch.rand(nx * ny * nz, device="cuda", generator=g) + 3.0
    cap = int(ray_off[-1]) * (nz - 1)
    out = frechet_rows(kernel, nx, ny, nz, kmax, *dev, cap=cap)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for _ in range(steps):
        out = frechet_rows(kernel, nx, ny, nz, kmax, *dev, cap=cap)
    ev[1].record()
    for _ in range(steps):
        grad = frechet_grad(kernel, nx, ny, nz, kmax, *dev, resid, vs)
    ev[2].record()
    torch.cuda.synchronize()
    ms_rows = ev[0].elapsed_time(ev[1]) / steps
    ms_grad = ev[1].elapsed_time(ev[2]) / steps
    entries = int(ray_off[-1]) * (nz - 1)
    nnz = int(out[1].numel())
    # algorithmic bytes: the count and fill passes each read the kernel value
    # (8 B) of every entry; the fill writes column and value of every nonzero
    # (8 B); the rays' cells and weights (8 B per cell) are read per depth
    rows_bytes = entries * 8 * 2 + nnz * 8 + int(ray_off[-1]) * 8 * 2
    grad_bytes = entries * 8 + int(ray_off[-1]) * 8
    return {"grid": [nx, ny, nz], "kmax": kmax, "rows": nrows, "ray_cells": int(ray_off[-1]),
            "entries": entries, "nnz": nnz, "rows_ms": ms_rows, "grad_ms": ms_grad,
            "rows_per_s": nrows / (ms_rows * 1e-3),
            "rows_GBps_algorithmic": rows_bytes / (ms_rows * 1e-3) / 1e9,
            "grad_GBps_algorithmic": grad_bytes / (ms_grad * 1e-3) / 1e9,
            "grad_norm": float(grad.norm().item())}


def main(argv=None):
    """python -m jdsurfg_amd.frechet [--rows N]: Fréchet rows and gradient of
    synthetic rays over a grid256-sized grid; prints one JSON report."""
    import argparse
    ap = argparse.ArgumentParser(prog="python -m jdsurfg_amd.frechet")
    ap.add_argument("--rows", type=int, default=20000)
    ap.add_argument("--kmax", type=int, default=120)
    a = ap.parse_args(argv)
    t0 = time.time()
    rep = bench(nrows=a.rows, kmax=a.kmax)
    rep["wall_s"] = time.time() - t0
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
