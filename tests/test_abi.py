"""The C-ABI boundary without a GPU: both libraries load, export every
function that include/*.h declares, and the ctypes mirrors of the
reference's structs have the reference's layout (src/polesitter.h:304-322)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT
from polesitter_amd import api

INCLUDE = os.path.join(ROOT, "include")


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = text.split("#ifdef POLESITTER_IMPLEMENTATION")[0]
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|size_t|void\s*\*|ps_result_t|struct\s+ps_hip_engine\s*\*)"
                       r"\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_hip_library_exports_everything_declared():
    L = api.hip_lib()
    names = declared("polesitter_hip.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def declared_arity(header):
    """parameter count of every function the header declares"""
    text = open(os.path.join(INCLUDE, header)).read()
    text = text.split("#ifdef POLESITTER_IMPLEMENTATION")[0]
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name in declared(header):
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", text)
        params = m.group(1).strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_signatures_match_the_header():
    # every ps_hip_* binding of the Python mirror passes as many arguments
    # as the C declaration takes
    L = api.hip_lib()
    arity = declared_arity("polesitter_hip.h")
    bound = 0
    for name, k in arity.items():
        f = getattr(L, name)
        if f.argtypes is None:
            continue
        bound += 1
        assert len(f.argtypes) == k, (name, len(f.argtypes), k)
    assert bound >= 20


def test_host_library_exports_the_reference_api():
    L = api.host_lib()
    names = declared("polesitter.h")
    for ref in ("ps_init", "ps_calc_forces", "ps_prepare_particles", "ps_destroy"):
        assert ref in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layout_matches_reference():
    # field order of H:304-310 and H:312-322 (LP64)
    assert [f[0] for f in api.PsConfig._fields_] == ["buff", "buff_size", "max_particles",
                                                     "theta", "thrd_cnt"]
    assert C.sizeof(api.PsConfig) == 40 and api.PsConfig.theta.offset == 24
    assert [f[0] for f in api.PsParticleArrs._fields_] == ["x", "y", "z", "mass", "fx", "fy",
                                                           "fz", "id", "cnt"]
    assert C.sizeof(api.PsParticleArrs) == 72
    # the extension keeps ps_config_t whole as its first member
    assert api.PsConfigEx.base.offset == 0 and api.PsConfigEx.dev_buff.offset == 40
    assert C.sizeof(api.PsConfigEx) == 56


def test_extended_config_leaves_reference_initialisers_warning_free(tmp_path):
    # appending fields to ps_config_t itself would make the reference's
    # 5-value initialisers miss fields, which -Wextra reports; the extension
    # is a separate struct, so both forms compile with -Werror
    src = tmp_path / "x.c"
    src.write_text('#define POLESITTER_IMPLEMENTATION\n#include "polesitter.h"\n'
                   'int main(void){ static char b[1<<20]; ps_context_t* c = 0;\n'
                   ' ps_config_t cfg = {b, sizeof b, 4, 2.0F, 4};\n'
                   ' ps_config_ex_t ex = {{b, sizeof b, 4, 2.0F, 4}, 0, 0};\n'
                   ' (void)cfg; return ps_init_ex(&c, 0) == PS_EINVAL && ex.dev_buff == 0 ? 0 : 1; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        f"-I{INCLUDE}", "-c", str(src), "-o", str(tmp_path / "x.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_header_compiles_as_c99_with_positional_initialisers(tmp_path):
    # the reference's callers initialise ps_config_t positionally
    # (tests/main.c:82): the drop-in header must accept that unchanged
    src = tmp_path / "t.c"
    src.write_text('#define PS_MULTITHREADING\n#define POLESITTER_IMPLEMENTATION\n'
                   '#include "polesitter.h"\n'
                   'int main(void){ static char b[1<<20]; ps_context_t* c = 0;\n'
                   ' ps_config_t cfg = {b, sizeof b, 4, 3.4641F, 4};\n'
                   ' ps_particle_arrs_t a = {0,0,0,0,0,0,0,0,0};\n'
                   ' (void)a; return ps_init(0, &cfg) == PS_EINVAL && c == 0 ? 0 : 1; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        f"-I{INCLUDE}", "-c", str(src), "-o", str(tmp_path / "t.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_invalid_arguments_without_device():
    # argument checks that return before any device work, as in H:1700,1843,1914
    L = api.host_lib()
    assert L.ps_init(None, None) == api.PS_EINVAL
    rc, ctx = api.ps_init(16, 4)  # smaller than the context itself
    assert rc == api.PS_EINVAL and ctx is None
    assert api.ps_destroy(None) == api.PS_EINVAL
    arrs = api.make_arrs(None, None, None, None, None, None, None, None, 0)
    rc, *_ = api.ps_prepare_particles(arrs, None)
    assert rc == api.PS_EINVAL
    assert L.ps_calc_forces(None, None, None, 0.0, 0.0, 0.0, 1.0) == api.PS_EINVAL


def test_failure_hook_arms_and_disarms_without_device():
    L = api.hip_lib()
    assert L.ps_hip_debug_fail(7, 0) == api.PS_EINVAL  # unknown kind
    for what in (api.FAIL_ALLOC, api.FAIL_WAIT):
        api.debug_fail(what, 3)
        api.debug_fail(what, -1)
    api.debug_fail(api.FAIL_OFF)


def _has_device():
    try:
        api.device_name()
        return True
    except api.PsError:
        return False


def test_init_without_device_is_a_device_error():
    # ps_init reaches the runtime, which has no device here: PS_EDEVICE, and
    # no context comes back
    if _has_device():
        pytest.skip("a device is present")
    rc, ctx = api.ps_init(api.buff_size_for(1000), 1000, 2.0)
    assert rc == api.PS_EDEVICE and ctx is None


def test_dropin_program_is_built():
    assert os.access(os.path.join(api.LIBDIR, "dropin_tick"), os.X_OK)


def integration_c_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"^```c\n(.*?)^```", text, flags=re.S | re.M)


def test_integration_examples_compile_as_c99(tmp_path):
    # every C example a maintainer would copy from INTEGRATION.md is a whole
    # translation unit that compiles against the drop-in headers
    blocks = integration_c_blocks()
    assert len(blocks) >= 2
    for k, code in enumerate(blocks):
        src = tmp_path / f"example{k}.c"
        src.write_text(code)
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                            f"-I{INCLUDE}", "-c", str(src), "-o", str(tmp_path / f"e{k}.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, (k, r.stderr)


def test_integration_config_initialiser_uses_reference_field_order():
    # ps_config_t is {buff, buff_size, max_particles, theta, thrd_cnt}
    # (H:304-310); every positional initialiser in the documentation starts
    # with the buffer
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    inits = re.findall(r"ps_config_t\s+\w+\s*=\s*\{([^}]*)\}", text)
    assert inits
    for body in inits:
        first = body.split(",")[0].strip()
        assert "buff" in first or first.startswith("b"), body
