"""The C-ABI library loads and exports every entry point include/asam_sound.h and
include/asam_ensemble.h declare,
and the ctypes mirrors of the header structs have the C layout (no GPU calls)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(sound_mod):
    L = sound_mod.load()
    names = sound_mod.exported_names_from_header()
    assert len(names) >= 45 and "asam_ens_solve_device" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", sound_mod.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_struct_layouts_match_header(sound_mod):
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "asam_ensemble.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(asam_sound_grid),
         sizeof(asam_sound_control), sizeof(asam_sound_info), sizeof(asam_sound_tuning),
         offsetof(asam_sound_grid, beta0_dt), offsetof(asam_sound_info, ms),
         sizeof(asam_ens_tuning), sizeof(asam_ens_stats));
  return 0;
}
'''
    import ctypes as C
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    S = sound_mod
    from asam_amd import ensemble as E
    assert vals == [C.sizeof(S.Grid), C.sizeof(S.Control), C.sizeof(S.Info),
                    C.sizeof(S.Tuning), S.Grid.beta0_dt.offset, S.Info.ms.offset,
                    C.sizeof(E.EnsTuning), C.sizeof(E.EnsStats)]


def test_no_cpu_fallback_in_product(sound_mod):
    """The product path never imports the oracle: the host mirror fails loudly without
    the HIP library."""
    for f in ("sound.py", "ensemble.py", "rosenbrock.py", "workload.py", "weights.py"):
        src = open(os.path.join(ROOT, "asam_amd", f)).read()
        assert "import oracle" not in src and "from oracle" not in src, f
        assert "pyoracle" not in src and "liborc" not in src, f
    with pytest.raises(ImportError):
        sound_mod.load.__wrapped__ if hasattr(sound_mod.load, "__wrapped__") else None
        import importlib
        m = importlib.import_module("asam_amd.sound")
        saved = m._lib
        m._lib = None
        try:
            m.load(os.path.join(ROOT, "does_not_exist.so"))
        finally:
            m._lib = saved


def test_method_constants_match_header(sound_mod):
    """The Python and Fortran method codes equal the C header's ASAM_METHOD_* values."""
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "asam_sound.h")).read()
    c_codes = {m: int(v) for m, v in re.findall(r"#define ASAM_METHOD_(\w+) (\d+)", hdr)}
    assert c_codes == sound_mod.METHODS
    f90 = open(os.path.join(root, "asam_amd", "fortran", "asam_sound_iface.f90")).read()
    f_codes = {m: int(v) for m, v in
               re.findall(r"ASAM_METHOD_(\w+) = (\d+)", f90)}
    assert f_codes == c_codes
