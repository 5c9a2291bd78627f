"""The C-ABI library loads, exports every symbol include/emc.h declares, and
reports errors with status codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import REPO, cfg1_deck, emcgpu


def header_functions():
    src = open(os.path.join(REPO, "include", "emc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(emc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(emcgpu.LIB_PATH)
    names = header_functions()
    assert len(names) >= 24
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    sig = emcgpu.abi.declare(emcgpu.load_library())
    assert set(header_functions()) <= set(sig), set(header_functions()) - set(sig)


def test_host_only_context_and_views():
    eng = emcgpu.Engine(cfg1_deck(), device=-1)
    m = eng.mesh
    assert (m.cnx, m.cny, m.cnz) == (20, 40, 12)
    assert m.nnode == 21 * 41 * 13
    b = eng.band
    assert b.nt == 55320 and b.nk == 12619 and list(b.nband) == [4, 3, 1]
    n = eng.particles_init()
    p = eng.download()
    assert p.n == n and n > 20000
    assert set(np.unique(p.ptype)) == {0, 1}
    # the flight refuses to run without a GPU instead of falling back to the CPU
    with pytest.raises(emcgpu.EmcError) as e:
        eng.fly_step(0)
    assert e.value.code == emcgpu.abi.EMC_ERR_NOGPU
    eng.close()


def test_errors_have_codes(tmp_path):
    with pytest.raises(emcgpu.EmcError) as e:
        emcgpu.Engine(str(tmp_path / "missing.txt"), device=-1)
    assert e.value.code == emcgpu.abi.EMC_ERR_IO
    # a deck coordinate off the grid lines is rejected (distance2index, src/mcmodel.cpp:4168-4176)
    d = tmp_path / "bad"
    emcgpu.write_deck(str(d))
    ldg = (d / "ldg.txt").read_text().replace("donor 0 10 -20 -7 0 5 1e26", "donor 0 10 -20 -7.5 0 5 1e26")
    (d / "ldg.txt").write_text(ldg)
    with pytest.raises(emcgpu.EmcError) as e:
        emcgpu.Engine(str(d / "input.txt"), device=-1)
    assert e.value.code == emcgpu.abi.EMC_ERR_DECK and "not aligned" in str(e.value)
    with pytest.raises(emcgpu.EmcError) as e:
        emcgpu.Engine(cfg1_deck(), device=-1, table_dir=str(tmp_path))
    assert e.value.code == emcgpu.abi.EMC_ERR_IO


def test_upload_validates_indices():
    eng = emcgpu.Engine(cfg1_deck(), device=-1)
    eng.particles_init()
    p = eng.download()
    p.itet[3] = 10 ** 6
    with pytest.raises(emcgpu.EmcError) as e:
        eng.upload(p)
    assert e.value.code == emcgpu.abi.EMC_ERR_ARG
    eng.close()


def test_checkpoint_and_refresh_need_a_gpu(tmp_path):
    eng = emcgpu.Engine(cfg1_deck(), device=-1)
    try:
        lib = eng.lib
        assert lib.emc_checkpoint_save(eng.ctx, str(tmp_path / "x").encode()) == emcgpu.abi.EMC_ERR_NOGPU
        assert lib.emc_checkpoint_load(eng.ctx, str(tmp_path / "x").encode()) == emcgpu.abi.EMC_ERR_NOGPU
        b, a = C.c_int64(0), C.c_int64(0)
        assert lib.emc_multiple_refresh(eng.ctx, 9, C.byref(b), C.byref(a)) == emcgpu.abi.EMC_ERR_NOGPU
    finally:
        eng.close()


EMC_RUN = os.path.join(REPO, "c-mc_electron_amd", "bin", "emc_run")


def test_emc_run_executable_finds_its_tables(tmp_path):
    """`emc_run input.txt` reads the deck from the working directory and the
    tables from dirname(exe)/../input: without a GPU it gets as far as the run
    and stops there with the no-GPU status (a missing table would stop it at
    emc_create with an I/O error first)."""
    import subprocess
    assert os.access(EMC_RUN, os.X_OK)
    emcgpu.write_deck(str(tmp_path), refine=1, electrons=2000, holes=200, vd=0.5, vg=0.7, total_step=2,
                      stat_step=1)
    r = subprocess.run([EMC_RUN, "input.txt", "--device", "-1"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1
    assert "a run needs a GPU" in r.stderr, r.stderr
    assert os.path.isdir(tmp_path / "data")
    r = subprocess.run([EMC_RUN], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
