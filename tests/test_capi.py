"""The C ABI library (no GPU needed): it loads, exports every function
include/k3dom.h declares, its structs match the header byte for byte, and
the host-only entry points (parameter loading, argument validation) work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "k3dom.h")
REF = "/root/reference"


@pytest.fixture(scope="module")
def lib():
    from k3dom_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from k3dom_amd import build
        build.build_lib()
    return _lib.load()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(k3dom_\w+)\s*\(", text, flags=re.M)))


def test_exports_every_declared_function(lib):
    names = declared_functions()
    assert len(names) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "k3dom_amd", "libk3dom.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    from k3dom_amd._lib import SIGNATURES
    assert set(SIGNATURES) == set(names)


def test_struct_layout_matches_header(tmp_path, lib):
    from k3dom_amd._lib import Params, Cell, Info, Lidar
    src = tmp_path / "lay.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "k3dom.h"', "int main(void){"]
    for cname, py in [("k3dom_params", Params), ("k3dom_cell", Cell), ("k3dom_info", Info), ("k3dom_lidar", Lidar)]:
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    for cname, py in [("k3dom_params", Params), ("k3dom_cell", Cell), ("k3dom_info", Info), ("k3dom_lidar", Lidar)]:
        assert int(got[cname]) == C.sizeof(py)
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
    assert C.sizeof(Cell) == 68


def f32(x):
    return np.float32(x)


def test_params_default_are_node_defaults(lib):
    from k3dom_amd.dom import params_default
    p = params_default()
    r = f32(0.1)
    assert p.resolution == r and p.size == f32(10.0) and p.size_z == f32(5.0) and p.mass_scale == f32(3.0)
    assert p.particle_count == 30_000_000 and p.new_born_particle_count == 3_000_000
    assert p.stddev_process_noise_position == f32(f32(0.1) / r)
    assert p.stddev_process_noise_velocity == f32(f32(1.0) / r)
    assert p.init_max_velocity == f32(f32(5.0) / r) and p.particle_min_vel == f32(f32(1.0) / r)
    assert p.sigma == 1.0 and p.ls == 1.0 and p.gamma == f32(0.9)
    assert p.prior_free == p.prior_static == p.prior_dynamic == 1.0 and p.prior_occ == 0.0
    assert p.map_shift_thresh == int(f32(1.0) / r)
    assert p.seed == 123456


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_params_yaml_conversion(tmp_path, lib):
    from k3dom_amd.dom import params_from_yaml
    m = _write(tmp_path, "m.yaml", "free_thresh: free_thresh\nparticle_count: 1000000 # one million\n"
                                  "new_born_particle_count: 100000\nstddev_process_noise_position: 0.01\n"
                                  "stddev_process_noise_velocity: 0.1\ninit_max_velocity: 5.0\n"
                                  "particle_min_vel: 0.1\nsigma: 0.1\nls: 0.5\nmass_scale: 2.5\n"
                                  "prior_free: 0.001\nprior_static: 0.001\nprior_dynamic: 0.001\ngamma: 0.99\n"
                                  "map_shift_thresh: 1\nresolution: 0.5\n")
    d = _write(tmp_path, "d.yaml", "resolution: 0.2\nsize: 40.0\nsize_z: 5.0\nsensor_off_x: 5.0\n"
                                  "sensor_off_y: -4.7\npose_imu_offset: 0\n")
    p = params_from_yaml(m, d)
    r = f32(0.2)
    assert p.resolution == r and p.size == 40.0 and p.size_z == 5.0
    assert p.sensor_off_x == 5.0 and p.sensor_off_y == f32(-4.7) and p.sensor_off_z == 0.0
    assert p.particle_count == 1_000_000 and p.new_born_particle_count == 100_000
    assert p.stddev_process_noise_position == f32(f32(0.01) / r)
    assert p.stddev_process_noise_velocity == f32(f32(0.1) / r)
    assert p.init_max_velocity == f32(f32(5.0) / r) and p.particle_min_vel == f32(f32(0.1) / r)
    assert p.map_shift_thresh == int(f32(1.0) / r)
    assert p.ls == 0.5 and p.sigma == f32(0.1) and p.gamma == f32(0.99) and p.mass_scale == 2.5


def test_params_yaml_missing_file(lib):
    from k3dom_amd.dom import params_from_yaml
    from k3dom_amd._lib import K3domError
    with pytest.raises(K3domError):
        params_from_yaml("/nonexistent/k3dom.yaml")


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "config")), reason="the reference's config files are not on this machine")
def test_params_from_reference_yaml(lib):
    from k3dom_amd.dom import params_from_yaml
    for ds, res, size, size_z, off in [("complex", 0.2, 40.0, 5.0, (5.0, -4.7, 0.0)),
                                      ("simple", 0.2, 21.0, 3.0, (-6.0, 0.0, -0.5))]:
        p = params_from_yaml(f"{REF}/config/k3dom.yaml", f"{REF}/config/{ds}.yaml")
        r = f32(res)
        assert p.resolution == r and p.size == f32(size) and p.size_z == f32(size_z)
        assert (p.sensor_off_x, p.sensor_off_y, p.sensor_off_z) == tuple(f32(o) for o in off)
        assert p.particle_count == 1_000_000 and p.new_born_particle_count == 100_000
        assert p.sigma == f32(0.1) and p.ls == f32(0.5) and p.mass_scale == f32(2.5) and p.gamma == f32(0.99)
        assert p.birth_prob == f32(0.02) and p.persistence_prob == f32(0.99)
        assert p.particle_min_vel == f32(f32(0.1) / r) and p.init_max_velocity == f32(f32(5.0) / r)
        assert p.map_shift_thresh == int(f32(1.0) / r)
        assert p.prior_free == f32(0.001) and p.prior_occ == 0.0


def test_create_rejects_invalid_params_without_gpu(lib):
    from k3dom_amd.dom import params_default
    from k3dom_amd._lib import Params
    for field, val in [("resolution", 0.0), ("particle_count", 0), ("size", 0.01), ("ls", -1.0),
                       ("new_born_particle_count", -5)]:
        p = params_default()
        setattr(p, field, val)
        h = C.c_void_p()
        st = lib.k3dom_create(C.byref(p), C.byref(h))
        assert st == 1 and not h.value
        assert lib.k3dom_last_error(None)
    assert lib.k3dom_create(None, C.byref(C.c_void_p())) == 1
    assert lib.k3dom_destroy(None) == 0
