"""HDF5 field and distribution files in the reference layout, and restart.

The reference writes FP64 datasets of dimensions gSD::pLength() -- the last
extent padded to 2 (n/2 + 1) (Domain.h:53-57, MathVector.h:331-344) -- named
density, velocityX.., alpha, forceX.., vorticity (FieldList.h:65-86,
Writer.h:309-360)
and distribution0..Q-1 (Writer.h:385-440), and restarts by reading the
distribution back (Reader.h:106-158, Initialize.h:119-145).  The files are
read here by an independent ctypes reader (tests/h5util.py) and by h5dump."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import h5util
from metalbm_amd import (Algorithm, CollisionType, Config, ForceType, LatticeType, LBMError,
                         Precision)
from metalbm_amd.capi import LBM_E_INVALID, LBM_E_STATE

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not h5util.available(), reason="libhdf5 not present")]

AX = "XYZ"


def padded(shape):
    s = list(shape)
    s[-1] = 2 * (s[-1] // 2 + 1)
    return tuple(s)


def cfg(lat, shape, collision=CollisionType.BGK, **kw):
    nz = shape[2] if len(shape) == 3 else 1
    return Config(lattice=lat, collision=collision, nx=shape[0], ny=shape[1], nz=nz,
                  tau=0.5003 if collision == CollisionType.ELBM else 0.6, **kw)


CASES = [
    ("d2q9-bgk-kolmogorov", cfg(LatticeType.D2Q9, (12, 9), force=ForceType.Kolmogorov,
                                amplitude=(1e-4, 0, 0), wavelength=(12.0, 1, 1)), (12, 9)),
    ("d3q19-bgk", cfg(LatticeType.D3Q19, (6, 4, 5)), (6, 4, 5)),
    ("d3q27-elbm-f32", cfg(LatticeType.D3Q27, (4, 6, 8), CollisionType.ELBM,
                           storage=Precision.F32, compute=Precision.F32, shifted=True),
     (4, 6, 8)),
]


OUTPUTS = [(True, True, True), (False, False, False), (True, False, True)]


@pytest.mark.parametrize("out", OUTPUTS, ids=["alpha-force-vorticity", "none", "alpha-vorticity"])
@pytest.mark.parametrize("name,c,shape", CASES, ids=[x[0] for x in CASES])
def test_fields_file_layout(tmp_path, name, c, shape, out):
    """writeAlpha, writeForce and writeVorticity select the optional datasets
    (all on in every shipped input, Input_prod.in:68-72): alpha (2 for BGK),
    the force components (zero without a force) and the reference-scaled
    vorticity (Curl::executeSpace + normalize, Routine.h:129-132): one
    dataset `vorticity` in 2D, vorticityX/Y/Z in 3D (FieldList.h:40, 84)."""
    print(f"case hdf5 fields {name} {out}", flush=True)
    w_alpha, w_force, w_vort = out
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 5)
        with pytest.raises(LBMError) as e:
            a.write_fields(str(tmp_path), 0)
        assert e.value.code == LBM_E_STATE
        a.iterate(3, store=True)
        rho, u, al, F = a.store_fields()
        om = a.vorticity("reference")[0]
        if out != (True, True, True):
            a.set_field_output(*out)
        path = a.write_fields(str(tmp_path), 3)
    assert os.path.basename(path) == "field-3.h5"
    D = len(shape)
    names = ["density"] + [f"velocity{AX[d]}" for d in range(D)]
    expect = [rho] + [u[d] for d in range(D)]
    if c.collision == CollisionType.BGK:
        assert np.all(al == 2.0)  # the BGK collision's alpha (Collision.h:114-123)
    if c.force == ForceType.None_:
        assert not F.any()
    if w_alpha:
        names.append("alpha")
        expect.append(al)
    if w_force:
        names += [f"force{AX[d]}" for d in range(D)]
        expect += [F[d] for d in range(D)]
    vnames = ["vorticity"] if D == 2 else [f"vorticity{AX[d]}" for d in range(3)]
    if w_vort:
        names += vnames
        expect += [om[k] for k in range(len(vnames))]
    absent = {"alpha", *[f"force{AX[d]}" for d in range(D)], *vnames} - set(names)
    with h5util.H5File(path) as h:
        for n, v in zip(names, expect):
            arr = h.read(n)
            assert arr.shape == padded(shape), n
            body = arr[..., :shape[-1]]
            np.testing.assert_array_equal(body.reshape(-1), v, err_msg=n)
            assert not arr[..., shape[-1]:].any(), n  # padding is zero
        for n in absent:
            assert not h.exists(n), n
    # XDMF sidecar (Writer.h:451-560): attributes in writing order, dims
    # pLength[X] then sLength[1..]
    xmf = open(os.path.join(str(tmp_path), "field-3.xmf")).read()
    assert f'TopologyType="{D}DCoRectMesh" Dimensions="{" ".join(map(str, shape))}"' in xmf
    assert re.findall(r'Attribute Name="([^"]+)"', xmf) == names
    assert re.findall(r"field-3\.h5:/(\w+)", xmf) == names
    if shutil.which("h5dump") or os.path.exists("/opt/conda/bin/h5dump"):
        exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
        out = subprocess.run([exe, "-H", path], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stderr
        dims = "( " + ", ".join(map(str, padded(shape))) + " )"
        assert out.stdout.count("H5T_IEEE_F64LE") == len(names)
        assert out.stdout.count(f"SIMPLE {{ {dims} / {dims} }}") == len(names)


@pytest.mark.parametrize("name,c,shape", CASES, ids=[x[0] for x in CASES])
def test_distribution_round_trip_and_restart(tmp_path, name, c, shape):
    """Write at step 10, restart a fresh handle from the file, and continue:
    10 + 10 steps equal 20 uninterrupted steps bit for bit (the alpha warm
    start, which the reference does not save, is carried over by hand)."""
    print(f"case hdf5 restart {name}", flush=True)
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 9)
        a.iterate(10)
        f10 = fw = a.populations()
        if c.storage != Precision.F64:
            a.set_populations(f10)  # both runs continue from set(f10)
            fw = a.populations()
        path = a.write_distribution(str(tmp_path), 10)
        al10 = a.alpha() if c.collision == CollisionType.ELBM else None
        a.iterate(10)
        f20 = a.populations()
    assert os.path.basename(path) == "distribution-10.h5"
    with h5util.H5File(path) as h:
        for i in range(c.q):
            arr = h.read(f"distribution{i}")
            assert arr.shape == padded(shape)
            np.testing.assert_array_equal(arr[..., :shape[-1]].reshape(-1), fw[i])
            assert not arr[..., shape[-1]:].any()
        assert not h.exists(f"distribution{c.q}")
    with Algorithm(c) as b:
        b.read_distribution(str(tmp_path), 10)
        g10 = b.populations()
        if c.storage == Precision.F64:
            np.testing.assert_array_equal(g10, f10)
        else:
            # F32 storage holds f - w: the FP64 round trip is exact to the
            # storage's own rounding
            np.testing.assert_allclose(g10, fw, rtol=0, atol=2.0 ** -24)
            b.set_populations(f10)
        if al10 is not None:
            b.set_alpha(al10)
        b.iterate(10)
        g20 = b.populations()
    np.testing.assert_array_equal(g20, f20)


def test_read_errors(tmp_path):
    print("case hdf5 read errors", flush=True)
    c = CASES[0][1]
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 1)
        a.iterate(1, store=True)
        a.write_distribution(str(tmp_path), 1)
        fields = a.write_fields(str(tmp_path), 1)
        f0 = a.populations()
        with pytest.raises(LBMError) as e:
            a.read_distribution(str(tmp_path), 2)
        assert e.value.code == LBM_E_INVALID and "cannot open" in e.value.message
        with pytest.raises(LBMError) as e:
            a.read_distribution_hdf5(fields)
        assert "distribution0 not found" in e.value.message
        np.testing.assert_array_equal(a.populations(), f0)  # unchanged
        a.iterate(1)
    other = Config(**{**c.__dict__, "ny": c.ny + 2})
    with Algorithm(other) as b:
        with pytest.raises(LBMError) as e:
            b.read_distribution(str(tmp_path), 1)
        assert "padded extents" in e.value.message
