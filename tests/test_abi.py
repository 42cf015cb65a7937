"""The C ABI: libplume_hip.so loads, exports every entry point include/plume_hip.h declares, and the ctypes
structs match the header's layout (checked by compiling the header with gcc). No GPU calls."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "plume_hip.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(plm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_operator_surface():
    names = declared_functions()
    for required in ("plm_advect_maccormack", "plm_buoyancy", "plm_confinement", "plm_emit", "plm_turbulence",
                     "plm_divergence", "plm_rb_sweeps", "plm_vcycle", "plm_subtract_gradient", "plm_smoke_budget",
                     "plm_light_volume", "plm_raymarch", "plm_build_flags", "plm_max_abs3"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import plume_amd._lib as lib

    for name in declared_functions():
        assert hasattr(lib.LIB, name), name
    assert set(lib.SIGNATURES) <= set(declared_functions())


def test_every_declared_entry_point_has_a_binding():
    import plume_amd._lib as lib

    bound = set(lib.SIGNATURES) | {"plm_last_error", "plm_abi_version", "plm_norm_scratch_doubles",
                                    "plm_sweep_tile_shapes", "plm_adj_raymarch_record_stride", "plm_mg_batch_depth"}
    assert set(declared_functions()) == bound


LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "plume_hip.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
int main(void) {
    printf("plm_grid_info %zu\nplm_emitter %zu\nplm_view %zu\nplm_look %zu\n", sizeof(plm_grid_info),
           sizeof(plm_emitter), sizeof(plm_view), sizeof(plm_look));
    F(plm_grid_info, cell_m); F(plm_grid_info, origin_m); F(plm_grid_info, open_lo); F(plm_grid_info, open_hi);
    F(plm_grid_info, fill_hi); F(plm_grid_info, ghost_gap);
    F(plm_emitter, center_m); F(plm_emitter, radius_m); F(plm_emitter, half_size_m); F(plm_emitter, edge_softness);
    F(plm_emitter, velocity_mps); F(plm_emitter, outward); F(plm_emitter, noise_scale_m); F(plm_emitter, stream);
    F(plm_view, left_m); F(plm_view, perspective); F(plm_view, eye);
    F(plm_look, shade_layout); F(plm_look, shade_fill);
    return 0;
}
"""


def test_ctypes_structs_match_the_header(tmp_path):
    import plume_amd._lib as lib

    source = tmp_path / "probe.c"
    source.write_text(LAYOUT_PROBE)
    binary = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(HEADER.parent), str(source), "-o", str(binary)], check=True)
    lines = subprocess.run([str(binary)], check=True, capture_output=True, text=True).stdout.split("\n")
    facts = dict(line.rsplit(" ", 1) for line in lines if line)
    structs = {"plm_grid_info": lib.GridInfo, "plm_emitter": lib.EmitterParams, "plm_view": lib.ViewParams,
               "plm_look": lib.LookParams}
    for key, value in facts.items():
        if "." in key:
            struct, member = key.split(".")
            assert getattr(structs[struct], member).offset == int(value), key
        else:
            assert ctypes.sizeof(structs[key]) == int(value), key


def test_bad_arguments_are_rejected_without_a_launch():
    """Argument checks run on the host: a NULL array or a bad shape returns PLM_EINVAL and a message."""
    import plume_amd._lib as lib
    from plume_amd.sim.grid import grid_info

    info = grid_info((4, 4, 4), 0.1)
    rc = lib.LIB.plm_divergence(ctypes.byref(info), None, None, None, None, None)
    assert rc == -1
    assert b"NULL" in lib.LIB.plm_last_error()
    bad = grid_info((0, 4, 4), 0.1)
    assert lib.LIB.plm_zero_closed_faces(ctypes.byref(bad), None, None, None, None) == -1
    assert b"positive" in lib.LIB.plm_last_error()
    with pytest.raises(lib.PlumeHipError, match="cell_m"):
        lib.call("plm_divergence", ctypes.byref(grid_info((4, 4, 4), 0.0)), None, None, None, None, None)


def test_the_product_refuses_the_cpu_device():
    from plume_amd.sim.device import DeviceError, init_device

    with pytest.raises(DeviceError):
        init_device("cpu")


def test_no_alloc_budget_rejects_large_planes_on_the_host():
    """plm_smoke_budget_noalloc checks the plane against its static scratch before any HIP call."""
    import plume_amd._lib as lib

    shape = lib.int3((4, 1024, 512))
    rc = lib.LIB.plm_smoke_budget_noalloc(shape, ctypes.c_void_p(256), ctypes.c_void_p(256), ctypes.c_void_p(256),
                                          None, None)
    assert rc == -1
    assert b"static scratch" in lib.LIB.plm_last_error()
