"""Register allocation of the built gfx950 code: no kernel of liblime_hip.so spills VGPRs or
uses private (scratch) memory.  A spill or a scratch-resident array in one of these kernels
turns the FP64 state that should stay in registers into memory traffic.

Read on the CPU from the code objects' metadata (the NT_AMDGPU_METADATA note, MessagePack)
of every gfx950 entry of the offload bundles in the library's .hip_fatbin section."""
import os
import struct

import pytest

from lime_amd import _lib

BUNDLE_MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
NT_AMDGPU_METADATA = 32
SHT_NOTE = 7


def _sections(elf):
    shoff = struct.unpack_from("<Q", elf, 0x28)[0]
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", elf, 0x3A)
    secs = [struct.unpack_from("<IIQQQQ", elf, shoff + i * shentsize) for i in range(shnum)]
    strtab = secs[shstrndx][4]

    def name(off):
        end = elf.index(b"\0", strtab + off)
        return elf[strtab + off:end].decode()
    return [(name(s[0]), s[1], s[4], s[5]) for s in secs]


def _code_objects(lib_bytes, arch="gfx950"):
    fat = [s for s in _sections(lib_bytes) if s[0] == ".hip_fatbin"]
    assert fat, "no .hip_fatbin section"
    _, _, off, size = fat[0]
    blob = lib_bytes[off:off + size]
    out, pos = [], 0
    while True:
        i = blob.find(BUNDLE_MAGIC, pos)
        if i < 0:
            return out
        count = struct.unpack_from("<Q", blob, i + len(BUNDLE_MAGIC))[0]
        p = i + len(BUNDLE_MAGIC) + 8
        for _ in range(count):
            eoff, esize, tlen = struct.unpack_from("<QQQ", blob, p)
            p += 24
            triple = blob[p:p + tlen].decode()
            p += tlen
            if triple.endswith(arch):
                out.append(blob[i + eoff:i + eoff + esize])
        pos = i + 1


def _kernels(co):
    msgpack = pytest.importorskip("msgpack")
    out = []
    for _, typ, off, size in _sections(co):
        if typ != SHT_NOTE:
            continue
        p = off
        while p < off + size:
            namesz, descsz, ntype = struct.unpack_from("<III", co, p)
            p += 12
            owner = co[p:p + namesz]
            p += (namesz + 3) & ~3
            desc = co[p:p + descsz]
            p += (descsz + 3) & ~3
            if ntype == NT_AMDGPU_METADATA and owner.startswith(b"AMDGPU"):
                out += msgpack.unpackb(desc, raw=False)["amdhsa.kernels"]
    return out


@pytest.fixture(scope="module")
def kernels():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblime_hip.so is not built (python -m lime_amd.build)")
    with open(_lib.LIB_PATH, "rb") as f:
        data = f.read()
    ks = [k for co in _code_objects(data) for k in _kernels(co)]
    assert ks, "no gfx950 kernels found"
    return ks


def test_every_kernel_is_spill_free_and_scratch_free(kernels):
    bad = [(k[".name"], k[".vgpr_spill_count"], k[".private_segment_fixed_size"])
           for k in kernels if k[".vgpr_spill_count"] or k[".private_segment_fixed_size"]]
    assert not bad, bad


def test_the_headline_and_path_kernels_are_present(kernels):
    names = {k[".name"] for k in kernels}
    # the bordered n = 65 resident kernel with records and the double buffer (variant 0)
    assert "_ZN4lime21lindblad_rk4_residentILi5ELi16ELi1ELi4ELb0ELb1ELb1ELb1EEEvNS_12ResidentArgsE" in names
    for stem in ("lindblad_stage_tiled", "lindblad_apply_tiled", "lindblad_left_tiled",
                 "lindblad_ragged_persistent", "adjoint_contract", "adjoint_seed", "obs_reduce"):
        assert any(stem in nm for nm in names), stem
    headline = [k for k in kernels if "ILi5ELi16ELi1ELi4ELb0ELb1ELb1ELb1E" in k[".name"]][0]
    assert headline[".vgpr_count"] + headline[".agpr_count"] <= 128      # 4 waves per SIMD
