"""Copy the reference's own SHCI test data into tests/golden/ as fixtures.

Run once in a container that has the reference mounted at /root/reference:

    python tests/golden/make_golden.py

It copies data files only (FCIDUMPs, the golden determinant list, trusted
energies) and extracts the numbers of the golden variational log into
golden.json.  Nothing under the reference is executed or imported.
"""
import json
import os
import re
import shutil
import struct

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

COPIES = {
    "h8_631g/FCIDUMP": "examples/h8_631g/FCIDUMP",
    "h8_631g/dets": "examples/h8_631g/dets",
    "restart/FCIDUMP": "tests/SHCI/restart/FCIDUMP",
    "n2_RAS/FCIDUMP": "tests/SHCI/n2_RAS/FCIDUMP",
    "o2/FCIDUMP": "tests/SHCI/o2_det/FCIDUMP",
    "h2o/FCIDUMP": "tests/SHCI/integrals/h2o_FCIDUMP_symm",
    "h2h2/FCIDUMP": "tests/SHCI/integrals/H2H2_FCIDUMP_nosym",
    "multislater_rhf/dets.bin": "tests/DQMC/multislater_rhf/dets.bin",
    "h2he_rdm/pyscf_1RDM.npy": "tests/SHCI/H2He_rdm/pyscf_1RDM.npy",
    "h2he_rdm/pyscf_2RDM.npy": "tests/SHCI/H2He_rdm/pyscf_2RDM.npy",
    "h2co_rdm/FCIDUMP": "tests/SHCI/integrals/h2co_FCIDUMP_nosymm",
    "h2co_rdm/pyscf_1RDM.npy": "tests/SHCI/h2co_rdm/pyscf_1RDM.npy",
    "h2co_rdm/pyscf_2RDM.npy": "tests/SHCI/h2co_rdm/pyscf_2RDM.npy",
    # (tests/SHCI/o2_stoc uses the o2_det FCIDUMP, byte for byte)
    "o2_stoc/trusted2RDM.txt": "tests/SHCI/o2_stoc/trusted2RDM.txt",
    "ueg/FCIDUMP": "tests/SHCI/integrals/UEG_FCIDUMP",
}
ENERGIES = {
    "restart": "tests/SHCI/restart/trusted_hci.e",
    "n2_RAS": "tests/SHCI/n2_RAS/trusted_hci.e",
    "o2_det_var_plus_pt": "tests/SHCI/o2_det/trusted_hci.e",
    "o2_det_direct": "tests/SHCI/o2_det_direct/trusted_hci.e",
    "fullrestart": "tests/SHCI/fullrestart/trusted_hci.e",
    "restart_trev": "tests/SHCI/restart_trev/trusted_hci.e",
    "o2_det_trev": "tests/SHCI/o2_det_trev/trusted_hci.e",
    "o2_det_trev_direct": "tests/SHCI/o2_det_trev_direct/trusted_hci.e",
}


def main():
    for dst, src in COPIES.items():
        d = os.path.join(HERE, dst)
        os.makedirs(os.path.dirname(d), exist_ok=True)
        shutil.copyfile(os.path.join(REF, src), d)
    gold = {"trusted_energy": {}}
    for k, src in ENERGIES.items():
        with open(os.path.join(REF, src), "rb") as f:
            gold["trusted_energy"][k] = struct.unpack("<d", f.read(8))[0]
    # golden variational table of examples/h8_631g/dice.out
    rows = []
    with open(os.path.join(REF, "examples/h8_631g/dice.out")) as f:
        text = f.read()
    for m in re.finditer(r"^\s+(\d+)\s+0\s+1\.00e-06\s+(\d+)\s+(-?\d+\.\d+)\s+(\d+)\s+[\d.]+$", text, re.M):
        rows.append({"iter": int(m.group(1)), "ndets": int(m.group(2)), "energy": float(m.group(3)),
                     "davidson": int(m.group(4))})
    final = re.search(r"^\s+0\s+(-?\d+\.\d+)\s+[\d.]+$", text.split("VARIATIONAL CALCULATION RESULT")[1], re.M)
    gold["h8_631g"] = {"iterations": rows, "final_energy": float(final.group(1)),
                       "input": open(os.path.join(REF, "examples/h8_631g/dice.dat")).read()}
    # inputs of the reference tests, kept as text data
    gold["restart_input3"] = open(os.path.join(REF, "tests/SHCI/restart/input3.dat")).read()
    # (tests/SHCI/fullrestart uses the restart test's FCIDUMP, byte for byte)
    gold["h2he_rdm_input"] = open(os.path.join(REF, "tests/SHCI/H2He_rdm/input.dat")).read()
    gold["h2co_rdm_input"] = open(os.path.join(REF, "tests/SHCI/h2co_rdm/test_input.dat")).read()
    gold["o2_stoc_input"] = open(os.path.join(REF, "tests/SHCI/o2_stoc/input.dat")).read()
    gold["spin1rdm_input"] = open(os.path.join(REF, "tests/SHCI/spin1rdm/input.dat")).read()
    # (tests/SHCI/restart_trev and the o2 time-reversal tests use the restart
    # and o2_det FCIDUMPs, byte for byte)
    for name in ("restart/input2", "fullrestart/input2", "fullrestart/input3", "restart_trev/input2",
                 "restart_trev/input3", "o2_det_trev/input", "o2_det_trev_direct/input"):
        gold[name.replace("/", "_")] = open(os.path.join(REF, "tests/SHCI/%s.dat" % name)).read()
    gold["n2_RAS_input"] = open(os.path.join(REF, "tests/SHCI/n2_RAS/input.dat")).read()
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(gold, f, indent=1)


if __name__ == "__main__":
    main()
