"""Every kernel in the built library allocates at least 32 VGPRs per wave (read from the
kernel descriptors of the gfx950 code objects embedded in the library; no GPU needed).

Waves of kernels with smaller allocations computed wrong values on MI355X while
long-running compute kernels shared their SIMDs (tests/test_gpu_concurrency.py runs that
case); each kernel's first statement, PGA_KERNEL_VGPR_FLOOR(), raises the allocation."""
import os
import struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "posit_gpu_amd", "lib", "libposit_gpu_amd.so")
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def code_objects(blob):
    """gfx950 ELF images of every offload bundle in the library."""
    pos = blob.find(MAGIC)
    while pos >= 0:
        count = struct.unpack_from("<Q", blob, pos + 24)[0]
        p = pos + 32
        for _ in range(count):
            off, size, tlen = struct.unpack_from("<QQQ", blob, p)
            triple = blob[p + 24:p + 24 + tlen].decode()
            p += 24 + tlen
            if "gfx950" in triple and size:
                yield blob[pos + off:pos + off + size]
        pos = blob.find(MAGIC, pos + 1)


def kernel_descriptors(elf):
    """(name, compute_pgm_rsrc1) of every kernel descriptor symbol (*.kd) in an ELF64 image."""
    shoff, = struct.unpack_from("<Q", elf, 0x28)
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", elf, 0x3A)
    sections = []
    for i in range(shnum):
        name, typ, _flags, addr, off, size, link, _info, _align, entsize = struct.unpack_from(
            "<IIQQQQIIQQ", elf, shoff + i * shentsize)
        sections.append(dict(name=name, type=typ, addr=addr, off=off, size=size, link=link, entsize=entsize))
    for sec in sections:
        if sec["type"] != 2:  # SHT_SYMTAB
            continue
        strtab = sections[sec["link"]]
        for k in range(sec["size"] // sec["entsize"]):
            st_name, _st_info, _st_other, st_shndx, st_value, _st_size = struct.unpack_from(
                "<IBBHQQ", elf, sec["off"] + k * sec["entsize"])
            end = elf.index(b"\0", strtab["off"] + st_name)
            sym = elf[strtab["off"] + st_name:end].decode()
            if not sym.endswith(".kd") or st_shndx >= len(sections):
                continue
            host = sections[st_shndx]
            kd = host["off"] + (st_value - host["addr"])
            yield sym[:-3], struct.unpack_from("<I", elf, kd + 48)[0]


def test_every_kernel_allocates_at_least_32_vgprs():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    blob = open(LIB, "rb").read()
    kernels = {}
    for elf in code_objects(blob):
        for name, rsrc1 in kernel_descriptors(elf):
            kernels[name] = ((rsrc1 & 0x3F) + 1) * 8  # granulated count, 8 VGPRs per granule
    assert len(kernels) > 20, sorted(kernels)
    small = {k: v for k, v in kernels.items() if v < 32}
    assert not small, small
