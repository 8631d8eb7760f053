"""ORACLE — test infrastructure only. numpy restatement of user_compute_Ei.m
(codes/GPU_implementation/seed/user_compute_Ei.m:40-57): RWG-tested incident
plane wave with the centroid rule. Angles in radians (the Matlab converts
degrees at :44; the Julia twin test_fields at compres_single_block_spheres_for_sweep2.jl:61
omits pi for phi, and is not followed)."""
import numpy as np


def plane_wave(vertex, edges, ln, cent, k, theta, phi, rot_eh=np.pi / 2):
    th, ph, rot = theta, phi, rot_eh
    k_i = -np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    e_i = np.cos(rot) * np.array([-np.sin(ph), np.cos(ph), 0.0]) + np.sin(rot) * np.array(
        [np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
    Tp = edges[0].astype(int) - 1
    Tm = edges[1].astype(int) - 1
    pp = np.exp(-1j * k * (k_i @ cent[:, Tp]))
    pm = np.exp(-1j * k * (k_i @ cent[:, Tm]))
    rp = cent[:, Tp] - vertex[:, edges[2].astype(int) - 1]
    rm = vertex[:, edges[3].astype(int) - 1] - cent[:, Tm]
    return ln * (np.sum((e_i[:, None] * pp) * rp, axis=0) + np.sum((e_i[:, None] * pm) * rm, axis=0)) / 2
