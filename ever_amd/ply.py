"""Scene I/O: PLY files of primitives to device tensors and back.

Reads binary little-endian PLY files with one ``vertex`` element of float properties, in either form the
reference handles:
* a 3D Gaussian Splatting file (``opacity`` logit, raw ``scale_i``): the viewer's loader
  (splinetracers/fast_ellipsoid_splinetracer/ply_file_loader.cc:196-293) applies EVER's activations
  (NormalizeSceneData, :164-193);
* a file written by the reference's save_ply (utils/math_util.py:152-202: ``density`` and ``scale_i`` already
  activated), read back as stored.
The header is parsed on the host; the payload goes to the device in one copy and one kernel
(ever_scene_from_records) turns the records into the tracer's arrays, optionally in Morton order of the means
(ever_morton_order, the loader's SortComponentsToMortonOrder, :72-148).

Properties are found by name, so their order in the file is free (the reference's C++ loader assumes the 3DGS
order); f_rest_i must be consecutive, as every writer of these files stores them.
"""
import ctypes
import os

import numpy as np
import torch

from . import _capi as C


def read_ply(path):
    """(property names, (N, P) float32 array view of the payload). Raises ValueError with the reference loader's
    messages (ply_file_loader.cc:205-238) for a malformed header."""
    data = bytearray(os.path.getsize(path))  # writable, so the payload view can back a tensor without a copy
    with open(path, "rb") as f:
        f.readinto(data)
    pos = 0

    def line():
        nonlocal pos
        end = data.find(b"\n", pos)
        if end < 0:
            raise ValueError("Expected string to contain a line ending, but found none.")
        s = data[pos:end].rstrip(b"\r").decode("ascii", "replace")
        pos = end + 1
        return s

    if line() != "ply":
        raise ValueError('Expected file to have "ply" header.')
    if line() != "format binary_little_endian 1.0":
        raise ValueError('Expected file to have "format binary_little_endian 1.0" header.')
    parts = line().split()
    while parts and parts[0] in ("comment", "obj_info"):
        parts = line().split()
    if len(parts) != 3 or parts[0] != "element" or parts[1] != "vertex" or not parts[2].isdigit():
        raise ValueError('Expected file to have "element vertex X" in the header, where X is the number of '
                         'components.')
    n = int(parts[2])
    names = []
    while True:
        s = line()
        if s == "end_header":
            break
        p = s.split()
        if p and p[0] in ("comment", "obj_info"):
            continue
        if len(p) != 3 or p[0] != "property":
            raise ValueError(f'Expected property line have format "property <type> <name>", got "{s}"')
        if p[1] not in ("float", "float32"):
            raise ValueError(f"property {p[2]}: only float properties are supported, got {p[1]}")
        names.append(p[2])
    need = n * len(names) * 4
    if len(data) - pos < need:
        raise ValueError(f"File data is truncated, expected at least {need} bytes, got {len(data) - pos}.")
    rec = np.frombuffer(data, dtype="<f4", count=n * len(names), offset=pos).reshape(n, len(names))
    return names, rec


def _layout(names, activation):
    col = {nm: i for i, nm in enumerate(names)}

    def need(nm):
        if nm not in col:
            raise ValueError(f"PLY file has no property {nm}")
        return col[nm]

    L = C.PlyLayout()
    L.num_columns = len(names)
    for a, ax in enumerate("xyz"):
        L.mean[a] = need(ax)
    for a in range(3):
        L.scale[a] = need(f"scale_{a}")
        L.f_dc[a] = need(f"f_dc_{a}")
    for a in range(4):
        L.rot[a] = need(f"rot_{a}")
    rest = sorted((int(nm[7:]), i) for nm, i in col.items() if nm.startswith("f_rest_") and nm[7:].isdigit())
    if len(rest) % 3:
        raise ValueError("f_rest properties must come in multiples of 3")
    if rest:
        if [k for k, _ in rest] != list(range(len(rest))) or [i for _, i in rest] != list(
                range(rest[0][1], rest[0][1] + len(rest))):
            raise ValueError("f_rest_0 .. f_rest_{M-1} must be consecutive properties")
        L.f_rest0 = rest[0][1]
    L.num_rest = len(rest) // 3
    if activation is None:
        activation = "3dgs" if "opacity" in col and "density" not in col else "none"
    if activation == "3dgs":
        L.activation = C.EVER_PLY_ACTIVATE_3DGS
        L.opacity = need("opacity")
    elif activation == "none":
        L.activation = C.EVER_PLY_ACTIVATE_NONE
        L.density = need("density")
    else:
        raise ValueError("activation must be None, '3dgs' or 'none'")
    return L


def morton_order(means):
    """(N,) int32 device tensor: the primitives sorted by the 63-bit Morton code of their means."""
    if not (isinstance(means, torch.Tensor) and means.is_cuda and means.dtype == torch.float32
            and means.is_contiguous() and means.dim() == 2 and means.size(1) == 3):
        raise RuntimeError("means must be a contiguous float32 (N,3) CUDA tensor")
    n = means.shape[0]
    L = C.lib()
    wsb = L.ever_morton_workspace_bytes(n)
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=means.device)
    order = torch.empty((n,), dtype=torch.int32, device=means.device)
    with torch.cuda.device(means.device):
        C.check(L.ever_morton_order(C.ptr(means), n, C.ptr(order), C.ptr(ws), wsb, C.stream_ptr()),
                "ever_morton_order")
    return order


def scene_from_records(names, records, device="cuda", activation=None, morton=False):
    """Device tensors (means, scales, quats, densities, features) from PLY records; see load_scene."""
    device = torch.device(device)
    L = _layout(names, activation)
    n = records.shape[0]
    rec = torch.from_numpy(np.ascontiguousarray(records, dtype=np.float32)).to(device)
    out = dict(means=torch.empty((n, 3), dtype=torch.float32, device=device),
               scales=torch.empty((n, 3), dtype=torch.float32, device=device),
               quats=torch.empty((n, 4), dtype=torch.float32, device=device),
               densities=torch.empty((n,), dtype=torch.float32, device=device),
               features=torch.empty((n, 1 + L.num_rest, 3), dtype=torch.float32, device=device))
    order = None
    if morton and n > 0:
        mx = [L.mean[a] for a in range(3)]
        order = morton_order(rec[:, mx].contiguous())
    with torch.cuda.device(device):
        C.check(C.lib().ever_scene_from_records(C.ptr(rec), n, ctypes.byref(L), C.ptr(order), C.ptr(out["means"]),
                                                C.ptr(out["scales"]), C.ptr(out["quats"]), C.ptr(out["densities"]),
                                                C.ptr(out["features"]), C.stream_ptr()),
                "ever_scene_from_records")
    if order is not None:
        out["order"] = order
    return out


def load_scene(path, device="cuda", activation=None, morton=False):
    """Primitives of a PLY file as device tensors: dict(means, scales, quats, densities, features[, order]).

    activation None picks "3dgs" for files with an ``opacity`` and no ``density`` property (the viewer's
    activations) and "none" otherwise (values as stored by save_ply). morton=True reorders the primitives by the
    Morton code of their means (``order`` holds the source index of each row)."""
    names, rec = read_ply(path)
    return scene_from_records(names, rec, device, activation, morton)


def save_ply(means, scales, quats, densities, features, path):
    """utils/math_util.py:152-202: vertex properties x y z nx ny nz f_dc_* f_rest_* density scale_* rot_*, normals
    zero, SH coefficients channel-major after the DC ones; binary little-endian float32."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    to = lambda t: (t.detach().cpu().float().numpy() if isinstance(t, torch.Tensor) else np.asarray(t, np.float32))
    xyz = to(means).reshape(-1, 3)
    n = xyz.shape[0]
    f = to(features).reshape(n, -1, 3)
    f_dc = f[:, :1].transpose(0, 2, 1).reshape(n, -1)
    f_rest = f[:, 1:].transpose(0, 2, 1).reshape(n, -1)
    sc = to(scales).reshape(n, -1)
    q = to(quats).reshape(n, -1)
    names = ["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(f.shape[2])]
    names += [f"f_rest_{i}" for i in range((f.shape[1] - 1) * f.shape[2])] + ["density"]
    names += [f"scale_{i}" for i in range(sc.shape[1])] + [f"rot_{i}" for i in range(q.shape[1])]
    body = np.concatenate([xyz, np.zeros_like(xyz), f_dc, f_rest, to(densities).reshape(n, 1), sc, q], 1)
    head = "ply\nformat binary_little_endian 1.0\nelement vertex {}\n{}end_header\n".format(
        n, "".join(f"property float {nm}\n" for nm in names))
    with open(path, "wb") as fh:
        fh.write(head.encode("ascii"))
        fh.write(np.ascontiguousarray(body, dtype="<f4").tobytes())
