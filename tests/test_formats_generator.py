"""CPU: the product's host-side generator, structured mesh builder and binary
readers/writers, pinned against the reference's ex05 mesh file and an
independent numpy generator.  Uses only host entry points of the library."""
import os

import numpy as np
import pytest

import synth_np

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EX05 = os.path.join(GOLD, "ex05_gpuline.mesh.data.bin")


def test_numpy_structured_rule_reproduces_ex05_bytes():
    raw = np.fromfile(EX05, dtype="<i4")
    img = synth_np.file_image(synth_np.structured(60, 64))
    assert raw.size == img.size and np.array_equal(raw, img)


def test_product_structured_mesh_equals_ex05(tbs):
    m = tbs.structured_mesh(60, 64)
    ref = tbs.read_mesh(EX05)
    assert m.sizes() == ref.sizes() == [9, 3840, 7556, 15360, 60, 3840]
    for f in ("epoint", "ef", "fc", "linesize", "linepoint", "lines", "lineface"):
        assert np.array_equal(getattr(m, f), getattr(ref, f)), f
    # reader shift 1 -> 0 (read_mesh_data.F90:40-64): boundary ids become <= -2
    assert ref.ef.min() == -249 and ref.lines[0] == 0 and ref.lineface[0] == -190


@pytest.mark.parametrize("NL,NK", [(1, 1), (1, 7), (5, 1), (13, 29), (360, 384)])
def test_product_mesh_equals_numpy_rule(tbs, NL, NK):
    m = tbs.structured_mesh(NL, NK)
    ref = synth_np.structured(NL, NK)
    for f in ref:
        assert np.array_equal(getattr(m, f), ref[f]), f


@pytest.mark.parametrize("seed,nvar,shift", [(1, 9, 20.0), (4, 9, 12.0), (77, 3, 5.5), (2**40 + 3, 1, 0.0)])
def test_product_generator_bitwise_equals_numpy(tbs, seed, nvar, shift):
    m = tbs.structured_mesh(7, 11, nvar)
    a = tbs.generate_host(seed, nvar, m.nelem, m.nintface, shift)
    b = synth_np.generate(seed, nvar, m.nelem, m.nintface, shift)
    for k in a:
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k


def test_generator_statistics(tbs):
    a = tbs.generate_host(1, 9, 3840, 7556, 20.0)
    D = a["jacD"].reshape(-1, 9, 9)
    diag = np.diagonal(D, axis1=1, axis2=2)
    assert diag.min() >= 19.0 and diag.max() <= 21.0
    assert np.abs(a["jacO1"]).max() <= 0.1 and np.abs(a["rhs"]).max() <= 1.0
    assert abs(a["rhs"].mean()) < 0.02


def test_mesh_and_jacobian_round_trip(tbs, tmp_path):
    m = tbs.structured_mesh(6, 9)
    p = tmp_path / "gpuline.mesh.data.bin"
    tbs.write_mesh(p, m)
    m2 = tbs.read_mesh(p)
    for f in ("epoint", "ef", "fc", "linesize", "linepoint", "lines", "lineface"):
        assert np.array_equal(getattr(m, f), getattr(m2, f))
    jac = tbs.generate_host(3, 9, m.nelem, m.nintface, 20.0)
    q = tmp_path / "gpuline.jacobian.data.bin"
    tbs.write_jacobian(q, 9, m.nelem, m.nintface, jac)
    nvar, nelem, nif, j2 = tbs.read_jacobian(q)
    assert (nvar, nelem, nif) == (9, m.nelem, m.nintface)
    for k in ("jacD", "jacO1", "jacO2", "rhs"):
        assert np.array_equal(jac[k], j2[k])
    assert not j2["U0"].any()
    assert os.path.getsize(q) == 12 + 8 * (81 * nelem + 2 * 81 * nif + 2 * 9 * nelem)


def test_truncated_and_missing_files_are_errors(tbs, tmp_path):
    raw = open(EX05, "rb").read()
    p = tmp_path / "short.bin"
    p.write_bytes(raw[:-4])
    with pytest.raises(tbs.TbsError) as ei:
        tbs.read_mesh(p)
    assert ei.value.code == -5
    with pytest.raises(tbs.TbsError):
        tbs.read_mesh(tmp_path / "nope.bin")
    with pytest.raises(tbs.TbsError):
        tbs.read_jacobian(p)


def test_byte_models(tbs):
    m = tbs.structured_mesh(1200, 1280)
    assert m.nintface == 3069520
    assert abs(tbs.model_bytes(m, 8) / 1e9 - 8.6214) < 5e-4
    assert abs(tbs.model_bytes(m, 4) / 1e9 - 7.1292) < 5e-4
    # src/main.cxx:149-175 per sweep: 3.551 (solve) + 0.553 (copy/add) + 5.815 (residual) GB
    assert abs(tbs.reference_model_bytes(m) / 1e9 - 9.9193) < 1e-3
