"""The C-ABI library loads, exports every symbol include/mpm_engine.h declares,
and the ctypes mirrors of its structs match the C layout. No device calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

from mpm_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpm_engine.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(mpm2?_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    syms = header_symbols()
    assert len(syms) >= 42
    for s in syms:
        assert hasattr(lib, s), s
    assert syms == set(L.SIGNATURES), syms ^ set(L.SIGNATURES)


def test_struct_layout_matches_c(tmp_path):
    structs = {"mpm_config": L.MpmConfig, "mpm_material": L.MpmMaterial, "mpm_particles": L.MpmParticles,
               "mpm_conservation": L.MpmConservation, "mpm_kernel_time": L.MpmKernelTime,
               "mpm_mlp_train_config": L.MpmMlpTrainConfig, "mpm_mlp_norm": L.MpmMlpNorm,
               "mpm_batch_options": L.MpmBatchOptions, "mpm_batch_tuning": L.MpmBatchTuning,
               "mpm_snapshot_stats": L.MpmSnapshotStats, "mpm2_particles": L.Mpm2Particles,
               "mpm2_settings": L.Mpm2Settings, "mpm2_stats": L.Mpm2Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpm_engine.h"', "int main(void){"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            cf = "lambda" if fname == "lambda_" else fname
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_invalid_arguments_rejected_before_any_device_call():
    lib = L.load()
    h = C.c_void_p()
    assert lib.mpm_create(None, C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    cfg = L.MpmConfig()
    cfg.lower[:] = [0, 0, 0]
    cfg.upper[:] = [1, 1, 1]
    cfg.cell_size = 0.0
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    cfg.cell_size = 8.0  # domain smaller than one cell
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    cfg.cell_size = 1 / 32
    cfg.wall_mode[0] = 7
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    cfg.wall_mode[0] = 0
    cfg.strict_domain = 1  # the reference caps h to [4, 16]
    assert lib.mpm_create(C.byref(cfg), C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_step(None, 1e-3, 1) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_destroy(None) == L.MPM_OK
    m = L.MpmMaterial()
    m.model = 12  # past the reference's MODELS table
    assert lib.mpm_return_map_batch(C.byref(m), 0, 0.0, None, None, None, None, None, None) == L.MPM_ERR_UNSUPPORTED


def defaults(model, E=1.0):
    m = L.MpmMaterial()
    assert L.load().mpm_material_defaults(model, E, C.byref(m)) == L.MPM_OK
    return m


def test_material_defaults_are_the_reference_records():
    m = defaults(9, 2.0)
    assert (m.model, m.E, m.mu, m.lambda_) == (9, 2.0, pytest.approx(.8), pytest.approx(.8))
    want = dict(cohesion=.025, slope=.6, dilatancy=.15, tension=.01, alpha=1.5, plateau=.015, gamma=.01,
                densify=.02, beta=2.0, epsD=.7)  # src/three/pressure-plasticity.js:2
    for k, v in want.items():
        assert getattr(m, "pr_" + k) == pytest.approx(v)
    assert (m.fracture, m.fr_onset, m.fr_failure) == (0, pytest.approx(.08), pytest.approx(.16))
    assert (m.yield_strain, m.friction, m.th_Thot, m.pl_Tm) == (pytest.approx(.025), 35.0, 180.0, 1000.0)
    from mpm_amd import Material
    c = Material(9, 2.0, .8, .8).to_c()  # the Python record fills the same defaults
    for name, _ in L.MpmMaterial._fields_:
        if name in ("yield_strain", "friction"):  # 0 there selects the same defaults on the device
            assert getattr(c, name) == 0
            continue
        assert getattr(c, name) == pytest.approx(getattr(m, name)), name


def test_material_check_follows_the_reference_parameter_rules():
    lib = L.load()
    for model in range(12):
        assert lib.mpm_material_check(C.byref(defaults(model))) == L.MPM_OK
    assert lib.mpm_material_check(C.byref(defaults(-1))) == L.MPM_ERR_UNSUPPORTED
    # tests/pressure-plasticity.test.mjs:51: each of these pressure records is rejected
    for bad in (dict(alpha=0), dict(epsD=0), dict(densify=0), dict(dilatancy=.8), dict(tension=.2),
                dict(cohesion=float("nan")), dict(beta=float("inf"))):
        for model in (8, 9):
            m = defaults(model)
            for k, v in bad.items():
                setattr(m, "pr_" + k, v)
            assert lib.mpm_material_check(C.byref(m)) == L.MPM_ERR_INVALID_ARGUMENT, (model, bad)
    m = defaults(0)
    m.pr_alpha = 0  # the pressure record is only checked for models 8 and 9
    assert lib.mpm_material_check(C.byref(m)) == L.MPM_OK
    for ok in (dict(alpha=.5), dict(alpha=2.0), dict(alpha=4.5 ** .5)):
        m = defaults(9)
        for k, v in ok.items():
            setattr(m, "pr_" + k, v)
        assert lib.mpm_material_check(C.byref(m)) == L.MPM_OK, ok
    # src/three/fracture.js:8: .005 <= onset < failure <= 2
    for onset, failure, rc in ((.04, .1, L.MPM_OK), (.004, .1, L.MPM_ERR_INVALID_ARGUMENT),
                               (.1, .1, L.MPM_ERR_INVALID_ARGUMENT), (.1, 2.5, L.MPM_ERR_INVALID_ARGUMENT),
                               (.005, 2.0, L.MPM_OK)):
        m = defaults(0)
        m.fracture, m.fr_onset, m.fr_failure = 1, onset, failure
        assert lib.mpm_material_check(C.byref(m)) == rc, (onset, failure)
    m = defaults(0)
    m.E = 0
    assert lib.mpm_material_check(C.byref(m)) == L.MPM_ERR_INVALID_ARGUMENT


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="not built"):
        L.load()


def test_surrogate_entry_points_validate_on_the_host():
    lib = L.load()
    m = L.MpmMaterial()
    m.model, m.E, m.th_A = 10, 2.0, 0.02
    c = L.MpmMlpTrainConfig(samples=1024, batch=256, epochs=1, seed=1, lr=1e-3, lr_final=1e-5, strain_sigma=0.02,
                            T_lo=300.0, T_hi=900.0, dt=1e-4, eps_max=0.0, near_yield_fraction=0.3)
    n = L.MpmMlpNorm()
    assert lib.mpm_mlp_default_norm(C.byref(m), C.byref(c), C.byref(n)) == L.MPM_OK
    assert n.in_scale == pytest.approx(0.5) and n.T_center == 600.0 and n.T_scale == pytest.approx(1 / 300)
    assert n.out_scale == pytest.approx(1.0) and n.eps_max == pytest.approx(0.08) and n.material == -1
    assert (n.T_lo, n.T_hi) == (300.0, 900.0) and n.min_yield_ratio == pytest.approx(1.0)
    # the fingerprint is E-normalised: a material with another E but the same
    # ratios matches, one with another yield ratio does not
    assert lib.mpm_mlp_check_material(C.byref(m), C.byref(n)) == L.MPM_OK
    m2 = L.MpmMaterial.from_buffer_copy(m)
    m2.E, m2.mu, m2.lambda_ = m.E * 1e5, m.mu * 1e5, m.lambda_ * 1e5
    assert lib.mpm_mlp_check_material(C.byref(m2), C.byref(n)) == L.MPM_OK
    m2.th_A = 0.03
    assert lib.mpm_mlp_check_material(C.byref(m2), C.byref(n)) == L.MPM_ERR_INVALID_ARGUMENT
    c.T_hi = 100.0
    assert lib.mpm_mlp_default_norm(C.byref(m), C.byref(c), C.byref(n)) == L.MPM_ERR_INVALID_ARGUMENT
    c.T_hi = 900.0
    P = (C.c_float * L.MLP_NPARAM)()
    assert lib.mpm_mlp_train(None, C.byref(c), C.byref(n), P, None) == L.MPM_ERR_INVALID_ARGUMENT
    c.batch = 4096  # larger than the sample count
    assert lib.mpm_mlp_train(C.byref(m), C.byref(c), C.byref(n), P, None) == L.MPM_ERR_INVALID_ARGUMENT
    m.model = 0  # no plastic return to learn
    assert lib.mpm_mlp_samples(C.byref(m), C.byref(c), C.byref(n), 0, 1, None, None) == L.MPM_ERR_UNSUPPORTED
    assert lib.mpm_surrogate_return_batch(C.byref(m), P, C.byref(n), 0, 0.0, None, None, None, None, None, None,
                                          None) == L.MPM_ERR_UNSUPPORTED
    m.model, m.th_H = 10, 0.5  # hardening: the return depends on the plastic history
    assert lib.mpm_surrogate_return_batch(C.byref(m), P, C.byref(n), 0, 0.0, None, None, None, None, None, None,
                                          None) == L.MPM_ERR_UNSUPPORTED
    assert lib.mpm_set_surrogate(None, 0, P, C.byref(n)) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_surrogate_counts(None, None, None, None, 0) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_mlp_loss_grad(None, 0, None, None, None, None) == L.MPM_ERR_INVALID_ARGUMENT


def test_batch_create_validates_on_the_host():
    lib = L.load()
    h = C.c_void_p()
    cfg = (L.MpmConfig * 2)()
    for c in cfg:
        c.lower[:] = [0, 0, 0]
        c.upper[:] = [1, 1, 1]
        c.cell_size = 1 / 32
    counts = (C.c_uint32 * 2)(100, 200)
    assert lib.mpm_batch_create(cfg, counts, 0, None, C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_batch_create(None, counts, 2, None, C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    cfg[1].cell_size = 0.0  # invalid scene geometry
    assert lib.mpm_batch_create(cfg, counts, 2, None, C.byref(h)) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_batch_destroy(None) == L.MPM_OK
    assert lib.mpm_batch_step(None, 1) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_batch_get_tuning(None, None) == L.MPM_ERR_INVALID_ARGUMENT
