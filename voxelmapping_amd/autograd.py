"""torch.autograd wrapper of the raycast and its hand-written adjoint.

``render(map, tsdf, xi, K, T0)`` returns the vertex and normal maps of the
map raycast at camToWorld = T0 exp(xi). Gradients flow to ``tsdf`` (a float32
tensor of n_block * 512 elements standing for the map's TSDF values; its
contents are not read, the map's own voxels are) and to ``xi`` (6 floats:
translation then rotation). The pose gradient is the one of a right-
multiplied twist at the current pose, which is d/dxi exactly at xi = 0.
"""
from __future__ import annotations

import numpy as np
import torch

from .mapping import DeviceMap, computeNormal


def hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]], dtype=np.float64)


def se3_exp(xi) -> np.ndarray:
    """4x4 exp of a twist (rho, phi), float64."""
    xi = np.asarray(xi, np.float64)
    rho, phi = xi[:3], xi[3:]
    th = np.linalg.norm(phi)
    Kx = hat(phi)
    if th < 1e-12:
        R = np.eye(3) + Kx
        V = np.eye(3) + 0.5 * Kx
    else:
        R = np.eye(3) + np.sin(th) / th * Kx + (1 - np.cos(th)) / th**2 * Kx @ Kx
        V = np.eye(3) + (1 - np.cos(th)) / th**2 * Kx + (th - np.sin(th)) / th**3 * Kx @ Kx
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ rho
    return T


class RaycastFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tsdf, xi, dmap: DeviceMap, K, T0, height: int, width: int):
        T = np.asarray(T0, np.float64) @ se3_exp(xi.detach().cpu().numpy())
        vmap = torch.zeros((height, width, 4), dtype=torch.float32, device=tsdf.device)
        dmap.raycast(vmap, K, T.astype(np.float32))
        nmap = computeNormal(vmap, torch.zeros_like(vmap))
        ctx.dmap, ctx.K, ctx.T = dmap, K, T.astype(np.float32)
        ctx.xi_device = xi.device
        ctx.save_for_backward(vmap)
        return vmap, nmap

    @staticmethod
    def backward(ctx, gvmap, gnmap):
        (vmap,) = ctx.saved_tensors
        dmap = ctx.dmap
        gt = torch.zeros(dmap.desc.n_block * 512, dtype=torch.float32, device=vmap.device)
        gp = torch.zeros(6, dtype=torch.float64, device=vmap.device)
        gv = None if gvmap is None else gvmap.contiguous()
        gn = None if gnmap is None else gnmap.contiguous()
        dmap.raycast_backward(vmap, ctx.K, ctx.T, gt, gp, grad_vertex=gv, grad_normal=gn, stats=False)
        return gt, gp.to(device=ctx.xi_device, dtype=torch.float32), None, None, None, None, None


def render(dmap: DeviceMap, tsdf, xi, K, T0, height: int, width: int):
    return RaycastFunction.apply(tsdf, xi, dmap, K, T0, height, width)
