"""Scenes of the family-A solver (01_acoustic_scattering.cpp) shared by the oracle and GPU
tests.  Geometry, density and probe positions are computed as 01 computes them, with the
C library's exp/log/sin/cos (Python's math module calls them), so that they are bit-equal."""
from __future__ import annotations

import math

import numpy as np

import nlos_engine as nlos


def config(x=(-8.0, 8.0), y=(-8.0, 8.0), z=(-0.2, 0.2), dx=0.05, c=1.0, tau=0.505,
           sponge=0.5, strength=1.0):
    """01's Solver members (01:58-66, 309) as an engine configuration."""
    cfg = nlos.default_config(x_min=x[0], x_max=x[1], y_min=y[0], y_max=y[1], z_min=z[0], z_max=z[1],
                              dx=dx, c_phys=c, tau=tau, sponge_width=sponge, sponge_strength=strength,
                              precision="f64")
    cfg.family = nlos.FAMILY_A
    return cfg


def dims(cfg):
    nx = int((cfg.x_max - cfg.x_min) / cfg.dx + 0.5)
    ny = int((cfg.y_max - cfg.y_min) / cfg.dx + 0.5)
    nz = max(3, int((cfg.z_max - cfg.z_min) / cfg.dx + 0.5))
    return nx, ny, nz


def cylinder_and_pulse(cfg, radius=0.5, centre=(4.0, 0.0), width=1.0, amp=0.01):
    """01's init_physics (01:166-205): solid where x^2 + y^2 < r^2; rho = 1 + amp *
    exp(-ln2 * d^2 / b^2) on fluid cells.  Arrays (nz, ny, nx)."""
    nx, ny, nz = dims(cfg)
    solid = np.zeros((nz, ny, nx), np.uint8)
    rho = np.ones((nz, ny, nx), np.float64)
    ln2 = math.log(2.0)
    plane_s = np.zeros((ny, nx), np.uint8)
    plane_r = np.ones((ny, nx), np.float64)
    for j in range(ny):
        y = cfg.y_min + (j + 0.5) * cfg.dx
        for i in range(nx):
            x = cfg.x_min + (i + 0.5) * cfg.dx
            if x * x + y * y < radius * radius:
                plane_s[j, i] = 1
            else:
                d2 = (x - centre[0]) * (x - centre[0]) + (y - centre[1]) * (y - centre[1])
                plane_r[j, i] = 1.0 + amp * math.exp(-ln2 * d2 / (width * width))
    solid[:] = plane_s
    rho[:] = plane_r
    return solid, rho


def ref01_probes():
    """01:157-163."""
    r = 135.0 * math.pi / 180.0
    return ["A", "B", "C"], [(0.0, 1.0, 0.0), (math.cos(r), math.sin(r), 0.0), (-1.0, 0.0, 0.0)]


def ref01_steps(cfg):
    """Records of 01's run loop (01:499-527): s = 0 .. (int)(11/dt) + 20."""
    dt = math.sqrt(1.0 / 3.0) * cfg.dx / cfg.c_phys
    return int(11.0 / dt) + 20


def radial_sponge_04(cfg):
    """04's init_physics sponge (04:176-208): beyond r = 10, ramp over 3, strength 0.5, with
    04's cell centres x = X_MIN + i dx + 0.5 dx.  Array (nz, ny, nx)."""
    nx, ny, nz = dims(cfg)
    sg = np.zeros((ny, nx))
    for j in range(ny):
        for i in range(nx):
            x = cfg.x_min + float(i) * cfg.dx + 0.5 * cfg.dx
            y = cfg.y_min + float(j) * cfg.dx + 0.5 * cfg.dx
            r = math.sqrt(x * x + y * y)
            if r > 10.0:
                ratio = min(1.0, (r - 10.0) / 3.0)
                sg[j, i] = 0.5 * ratio * ratio
    return np.broadcast_to(sg, (nz, ny, nx)).copy()


def cylinder_04(cfg, radius=4.0):
    """04's solid cylinder r < cyl_R along z (04:190-194), 04's cell centres."""
    nx, ny, nz = dims(cfg)
    plane = np.zeros((ny, nx), np.uint8)
    for j in range(ny):
        for i in range(nx):
            x = cfg.x_min + float(i) * cfg.dx + 0.5 * cfg.dx
            y = cfg.y_min + float(j) * cfg.dx + 0.5 * cfg.dx
            plane[j, i] = 1 if math.sqrt(x * x + y * y) < radius else 0
    return np.broadcast_to(plane, (nz, ny, nx)).copy()


def coarse04_config(dx=0.5, z=(-1.5, 1.5)):
    """04's +-13 m box and tau 0.51 (04:28-57) at a coarser dx."""
    return config(x=(-13.0, 13.0), y=(-13.0, 13.0), z=z, dx=dx, tau=0.51)
