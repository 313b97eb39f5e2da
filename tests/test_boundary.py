"""The C-ABI library on a machine without a GPU: it loads, exports every
entry point include/shci_gpu.h declares, and its host-side input handling
fails the way the reference does (SHCI/input.cpp:373-382) before any device
call."""
import ctypes
import os
import re

import pytest

from dice_amd import shci

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "shci_gpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shci_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = shci.library()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n


def _run_expect(tmp_path, text, status, needle):
    p = tmp_path / "input.dat"
    p.write_text(text)
    with pytest.raises(shci.ShciError) as e:
        shci.do_variational(str(p), str(tmp_path / "log"))
    assert e.value.status == status
    assert needle in str(e.value)


def test_unknown_keyword_fails_like_reference(tmp_path):
    _run_expect(tmp_path, "nocc 2\n0 1\nend\nschedule\n0 1e-3\nend\nmaxiter 2\nbogusword 3\n", 1,
                "cannot read option bogusword")


def test_maxiter_below_schedule(tmp_path):
    _run_expect(tmp_path, "nocc 2\n0 1\nend\nschedule\n0 1e-3\n5 1e-4\nend\nmaxiter 2\n", 1,
                "maxiter should be greater than last entry of sweep_iter")


def test_missing_nocc(tmp_path):
    _run_expect(tmp_path, "schedule\n0 1e-3\nend\nmaxiter 2\n", 1, "nocc keyword has to be included")


def test_nocc_count_mismatch(tmp_path):
    _run_expect(tmp_path, "nocc 3\n0 1\nend\nschedule\n0 1e-3\nend\nmaxiter 2\n", 1, "nocc: 3 neq 2")


def test_outside_path_keywords_refused(tmp_path):
    _run_expect(tmp_path, "nocc 2\n0 1\nend\nschedule\n0 1e-3\nend\nmaxiter 2\ndolcc\n", 5,
                "outside the variational path")
    _run_expect(tmp_path, "nocc 2\n0 1\nend\nschedule\n0 1e-3\nend\nmaxiter 2\nDoThreeRDM\n", 5,
                "outside the variational path")


def test_synthetic_fcidump_format(tmp_path):
    """The generator writes the reference writer's layout (lib/shcitools.c:165-234)."""
    from dice_amd import synth
    p = tmp_path / "FCIDUMP"
    synth.write_molecular_fcidump(str(p), 0, 6, 3, 3, 8)
    lines = p.read_text().splitlines()
    assert lines[0].startswith(" &FCI NORB=6 ,NELEC=6 ,MS2=0,")
    assert lines[3].strip() == "&END"
    first = lines[4]
    assert len(first.split()[0]) == 18 and first[:20].strip() == first.split()[0]
    assert lines[-1].split()[1:] == ["0", "0", "0", "0"]


def test_private_rand_sequence_equals_the_c_library(tmp_path):
    """The library draws the reference's rand() sequence (alias sampling of
    the semistochastic PT, random Davidson guesses) from a private copy of
    the C library's generator, so that other code in the process cannot
    advance it between srand(seed) and the draws: its outputs equal
    srand(seed); rand() for seeds including 0 and ones above 2^31."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cstdlib>
#include "host_io.h"
int main() {
  const unsigned seeds[] = {0u, 1u, 7u, 12345u, 2147483647u, 2147483648u, 4294967295u, 1764763200u};
  long bad = 0;
  for (unsigned s : seeds) {
    srand(s);
    shci::LibcRand g(s);
    for (int i = 0; i < 20000; i++)
      if (rand() != g.next()) bad++;
  }
  std::printf("%ld %d\n", bad, RAND_MAX == shci::LibcRand::kMax);
  return 0;
}
''')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "dice_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["0", "1"]


def test_in_memory_create_rejects_bad_arguments():
    """shci_create_from_integrals checks its arguments before any device call
    (SHCI_BAD_INPUT)."""
    import numpy as np
    h1 = np.eye(4)
    eri8 = np.zeros(55)  # npair = 10 -> 55 packed values
    for norb_nelec in ((4, 0), (4, 9)):
        with pytest.raises(shci.ShciError) as e:
            shci.System.from_integrals(h1, eri8, norb_nelec[1])
        assert e.value.status == 1, str(e.value)
        assert "nelec" in str(e.value)
