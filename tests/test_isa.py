"""The built gfx950 device code carries no packed-FP32 arithmetic.

librr_hip.so is compiled with the device feature ``-packed-fp32-ops``
(pytransit_amd/build.py).  This test takes every gfx950 code object out of the
library's ``.hip_fatbin`` section (one offload bundle per source file),
disassembles it with ROCm's llvm-objdump and checks that no v_pk_fma_f32,
v_pk_mul_f32 or v_pk_add_f32 is in it.  As a check that the disassembly covers
the whole library, every kernel family must be found and the TSModel contraction
must hold its MFMA instructions.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, 'pytransit_amd', 'librr_hip.so')
LLVM = '/opt/rocm/llvm/bin'
MAGIC = b'__CLANG_OFFLOAD_BUNDLE__'


def _tool(name):
    path = os.path.join(LLVM, name)
    if not os.path.exists(path):
        pytest.skip(f'{path} not found')
    return path


@pytest.fixture(scope='module')
def disassembly(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip('librr_hip.so is not built')
    objcopy, bundler, objdump = (_tool(n) for n in ('llvm-objcopy', 'clang-offload-bundler', 'llvm-objdump'))
    tmp = tmp_path_factory.mktemp('isa')
    fatbin = tmp / 'fatbin.bin'
    subprocess.run([objcopy, '-O', 'binary', '--only-section=.hip_fatbin', LIB, str(fatbin)], check=True)
    data = fatbin.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, 'no offload bundle in .hip_fatbin'
    texts = []
    for j, s in enumerate(starts):
        end = starts[j + 1] if j + 1 < len(starts) else len(data)
        part = tmp / f'bundle{j}.bin'
        part.write_bytes(data[s:end])
        co = tmp / f'bundle{j}.co'
        subprocess.run([bundler, '--unbundle', '--type=o', f'--input={part}',
                        '--targets=hipv4-amdgcn-amd-amdhsa--gfx950', f'--output={co}'], check=True)
        res = subprocess.run([objdump, '-d', str(co)], check=True, capture_output=True, text=True)
        texts.append(res.stdout)
    return '\n'.join(texts), len(starts)


def test_no_packed_fp32(disassembly):
    text, nbundles = disassembly
    from pytransit_amd.build import SOURCES
    assert nbundles == len(SOURCES)  # one offload bundle per HIP source
    bad = sorted(set(re.findall(r'\bv_pk_(?:fma|mul|add)_f32\b', text)))
    assert not bad, f'packed FP32 instructions in the device code: {bad}'


def test_disassembly_covers_every_kernel_family(disassembly):
    text, _ = disassembly
    for kern in ('rr_flux_kernel', 'rr_setup_kernel', 'ts_ldm_mfma_kernel', 'ts_flux_rows_kernel',
                 'rr_lnlike_kernel', 'ma_flux_kernel', 'ec_flux_kernel', 'op_flux_kernel'):
        assert kern in text, kern
    assert re.search(r'\bv_mfma_f32_16x16x4_?f32\b', text)
