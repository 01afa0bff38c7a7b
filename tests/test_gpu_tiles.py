"""The device tile builder against the host one: the fused kernel's tiles
(headers, blobs -- block tasks, items, halo connectivity, coordinates -- and
halo-node lists) must be the same bytes.  Both follow the same partition
(Morton box tree) and per-tile rules, so any difference is a builder bug.
Covers H8 and T10 blocks (perturbed), the importers' irregular T10 meshes,
every kernel family (FP64 stiffness with the bank-spread halo numbering,
mixed precision, the DOT forms' smaller tiles) and other tile parameters."""
import os
import sys

import numpy as np
import pytest

from feasm import _lib as L
from feasm import core
from feasm import finetools as F

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FAMILIES = [("elastic fp64", L.LIN_ELASTIC, L.FP64), ("elastic mixed", L.LIN_ELASTIC, L.MIXED),
            ("mass fp64", L.DOT, L.FP64), ("diffusion mixed", L.DIFFUSION_ISO, L.MIXED)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def h8(n, amp=0.15, seed=3):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
    return L.H8, conn, core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp / max(n), seed)


def t10(n, seed=5):
    conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, *n, amp_vertex=0.1 / max(n), amp_mid=0.03 / max(n), seed=seed)
    return L.T10, conn, xyz


def check(ctx, name, etype, conn, xyz, families=FAMILIES, params=None):
    mesh = core.DeviceMesh(ctx, etype, conn, xyz)
    mesh.set_dofs(3)
    mesh.build_pattern()
    if params:
        mesh.set_tile_params(*params)
    for fname, form, prec in families:
        say("tiles", name, fname, "params", params, "elements", conn.shape[0])
        nt, diff, ms = mesh.tiles_compare(form, prec)
        say("   ", nt, "tiles, device build", round(ms, 2), "ms, first difference", diff)
        assert nt >= 1
        assert diff == -1
    mesh.close()


@pytest.mark.parametrize("n", [(1, 1, 1), (3, 3, 2), (7, 5, 3), (13, 11, 9)])
def test_h8_blocks(ctx, n):
    check(ctx, "h8 %s" % (n,), *h8(n))


@pytest.mark.parametrize("n", [(1, 1, 1), (4, 3, 2), (9, 8, 7)])
def test_t10_blocks(ctx, n):
    check(ctx, "t10 %s" % (n,), *t10(n))


def test_unperturbed_grid_has_equal_morton_codes_split_in_the_middle(ctx):
    # a thin slab: one axis of the lattice collapses, many nodes share codes
    conn, xyz = core.gen_h8block(1.0, 1.0, 0.001, 12, 12, 1)
    check(ctx, "slab", L.H8, conn, xyz)


@pytest.mark.parametrize("name", ["cylinder", "le10"])
def test_irregular_meshes(ctx, name):
    if name == "cylinder":
        out = F.import_NASTRAN(os.path.join(HERE, "golden", "meshes", "cylinder.nas"), expectfixedformat=True)
    else:
        out = F.import_ABAQUS(os.path.join(HERE, "golden", "meshes", "problem_LE10NAFEMS_MST10.inp"))
    conn, xyz = out["fesets"][0].conn, out["fens"].xyz
    check(ctx, name, L.T10, conn, xyz)


@pytest.mark.parametrize("params", [(8, 0, 0), (16, 40, 0), (128, 0, 0), (0, 80, 512)])
def test_other_tile_parameters(ctx, params):
    check(ctx, "h8 (9,8,7)", *h8((9, 8, 7)), families=FAMILIES[:3], params=params)


def test_larger_meshes(ctx):
    check(ctx, "h8 (40,40,40)", *h8((40, 40, 40)), families=FAMILIES[:1])
    check(ctx, "t10 (16,16,16)", *t10((16, 16, 16)), families=FAMILIES[1:3])
