"""CPU checks of the boundary: layouts, exported symbols, error handling."""
import ctypes as C
import os
import re
import subprocess

import pytest

from webgpuraytracer_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def test_ctypes_layouts_match_header_sizes():
    for name, (struct, size) in A.SIZES.items():
        assert C.sizeof(struct) == size, name


def test_header_static_asserts_hold_in_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "pt_abi.h"\n#include "pt_scene.h"\n#include "pt_unet.h"\n#include "pt_batch.h"\nint main(void){return 0;}\n')
    subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-I" + INC, str(src)], check=True)
    subprocess.run(["g++", "-std=c++17", "-x", "c++", "-fsyntax-only", "-I" + INC, str(src)], check=True)


def _declared(header, prefix):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_libpt_exports_every_declared_symbol():
    lib = os.path.join(A.LIB_DIR, "libpt.so")
    names = _declared("pt_abi.h", "pt_")
    assert "pt_upload_scene" in names and "pt_trace" in names and "pt_denoise" in names
    missing = [n for n in names if n not in _exported(lib)]
    assert not missing, missing
    A.load_pt()  # every binding resolves


@pytest.mark.parametrize("header,must", [("pt_unet.h", ["pt_unet_train", "pt_unet_denoise"]),
                                         ("pt_batch.h", ["pt_batch_frame", "pt_batch_add"])])
def test_libpt_exports_the_other_headers(header, must):
    lib = os.path.join(A.LIB_DIR, "libpt.so")
    names = _declared(header, "pt_")
    assert all(m in names for m in must)
    missing = [n for n in names if n not in _exported(lib)]
    assert not missing, missing


def test_libptscene_exports_every_declared_symbol():
    lib = os.path.join(A.LIB_DIR, "libptscene.so")
    names = _declared("pt_scene.h", "ptgen_")
    missing = [n for n in names if n not in _exported(lib)]
    assert not missing, missing
    A.load_scene_lib()


def test_ctypes_table_covers_the_headers():
    declared = set(_declared("pt_abi.h", "pt_")) | set(_declared("pt_unet.h", "pt_")) | set(_declared("pt_batch.h", "pt_"))
    assert declared == set(A.PT_FUNCS)
    assert set(_declared("pt_scene.h", "ptgen_")) == set(A.PTGEN_FUNCS)


def test_null_arguments_are_errors_not_crashes():
    lib = A.load_pt()
    assert lib.pt_create(None, 0) == A.PT_ERR_ARG
    assert lib.pt_upload_scene(None, None) == A.PT_ERR_ARG
    assert lib.pt_trace(None, None, 0, 0, 1, None, None) == A.PT_ERR_ARG
    assert lib.pt_denoise(None, None, 0) == A.PT_ERR_ARG
    assert lib.pt_read_accum(None, None, 0) == A.PT_ERR_ARG
    assert lib.pt_write_accum(None, None, 0) == A.PT_ERR_ARG
    assert lib.pt_unet_init(None, 0) == A.PT_ERR_ARG
    assert lib.pt_unet_train(None, 1, 1, 1e-3, 0, None) == A.PT_ERR_ARG
    assert lib.pt_unet_denoise(None, 1.0) == A.PT_ERR_ARG
    assert lib.pt_batch_create(None, 0) == A.PT_ERR_ARG
    assert lib.pt_batch_add(None, None, None) == A.PT_ERR_ARG
    assert lib.pt_batch_frame(None, None, None, None, None) == A.PT_ERR_ARG
    lib.pt_batch_destroy(None)
    lib.pt_destroy(None)


def test_create_reports_missing_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = A.load_pt()
    h = C.c_void_p()
    assert lib.pt_create(C.byref(h), 0) == A.PT_ERR_NO_DEVICE


def test_renderer_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from webgpuraytracer_amd import Renderer
    with pytest.raises(A.PTError):
        Renderer(0)


def test_node_addon_loads_and_exports():
    addon = os.path.join(ROOT, "webgpuraytracer_amd", "js", "build", "pt_addon.node")
    if not os.path.exists(addon):
        pytest.skip("addon not built")
    js = ("const t=require('%s');"
          "const want=['create','uploadScene','updateObjects','resize','trace','gbuffer','denoise','readAccum',"
          "'writeAccum','readOutput','readGBuffer','readDenoised','lastError','setOption','unetInit','unetSetParams',"
          "'unetDenoise','unetSetPrecision','batchCreate','batchAdd','batchFrame','readSkyCdf'];"
          "for (const k of want) if (typeof t.addon[k] !== 'function') { console.log('missing '+k); process.exit(3); }"
          "const p=t.packParams({eye:[1,2,3],ray00:[0,0,0],ray10:[0,0,0],ray01:[0,0,0],ray11:[0,0,0],"
          "sampleNumber:5,width:7,height:9,seed:11,skyWidth:4,skyHeight:2,skyTotalLum:1.5,skyCdfIndex:3});"
          "if (p.byteLength!==112) process.exit(4);"
          "const dv=new DataView(p); if (dv.getUint32(12,true)!==5||dv.getUint32(76,true)!==11||"
          "dv.getUint32(104,true)!==3) process.exit(5);"
          "if (t.flagsWord({hasMeshes:true,hasSkybox:true,misSkybox:true})!==(32|512|1024)) process.exit(6);"
          "const w=t.loadNpyF32(t.PRETRAINED); if (w.length!==1021904||!(Math.abs(w[0])<10)) process.exit(7);"
          "console.log('ok');") % os.path.join(ROOT, "webgpuraytracer_amd", "js", "tracer_native.js")
    r = subprocess.run(["node", "-e", js], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout
