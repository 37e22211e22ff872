"""CPU checks of the drop-in boundary: exported symbols and host-side error paths (no GPU needed)."""
import os
import re
import subprocess

import numpy as np
import pytest

import ppiamd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "shim", "_build")


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def declared_c_functions(header):
    text = open(header).read()
    return set(re.findall(r"^\s*(?:ppi_status|int32_t|void|const char\*)\s+(ppi_\w+)\s*\(", text, re.M))


def test_engine_exports_every_declared_entry_point():
    decl = declared_c_functions(os.path.join(ROOT, "include", "ppi_engine.h"))
    assert {"ppi_engine_create", "ppi_engine_evaluate", "ppi_engine_wait"} <= decl
    syms = exported(ppiamd.ENGINE_LIB)
    missing = decl - syms
    assert not missing, f"declared but not exported: {missing}"
    assert set(ppiamd.ENGINE_SYMBOLS) <= syms


def test_engine_library_loads_without_device():
    L = ppiamd.lib()
    assert L.ppi_abi_version() == 1
    n = ppiamd.device_count()
    assert n >= 0


@pytest.mark.parametrize("variant", ["abs", "cost"])
def test_shim_exports_reference_signatures(variant):
    """The C++ (mangled) signatures of accelerator.h / sequential.h, both PROFILING variants."""
    plain = exported(os.path.join(SHIM, variant, "libppi_interpreter.so"))
    prof = exported(os.path.join(SHIM, variant, "libppi_interpreter_prof.so"))
    # void acc_interpret(Symbol*, float*, int*, float*, int, void(*)(Population*), int(*)(bool**), Population*,
    #                    int*, int*, int*, int, int, float)
    assert "_Z13acc_interpretP6SymbolPfPiS1_iPFvP10PopulationEPFiPPbES4_S2_S2_S2_iif" in plain
    assert "_Z18acc_interpret_initiPPcjjjPPfiiii" in plain
    assert "_Z13seq_interpretP6SymbolPfPiS1_iS2_S2_iif" in plain
    assert "_Z18seq_interpret_initjPPfii" in plain
    assert "_Z21seq_interpret_destroyv" in plain
    # PROFILING adds `unsigned long sum_size_gen` after size and the print functions.
    assert "_Z13acc_interpretP6SymbolPfPimS1_iPFvP10PopulationEPFiPPbES4_S2_S2_S2_iif" in prof
    assert "_Z13seq_interpretP6SymbolPfPimS1_iS2_S2_iif" in prof
    assert "_Z14acc_print_timeby" in prof and "_Z14seq_print_timeby" in prof


def write_case(path, data, pop, best_size=1, alpha=1e-6):
    with open(path, "wb") as f:
        np.array([data.shape[0], data.shape[1], pop.n, pop.stride, best_size], np.int32).tofile(f)
        np.array([alpha], np.float32).tofile(f)
        np.ascontiguousarray(data, np.float32).tofile(f)
        pop.phenotype.astype(np.int32).tofile(f)
        pop.ephemeral.astype(np.float32).tofile(f)
        pop.size.astype(np.int32).tofile(f)


@pytest.mark.parametrize("argv,expect", [(["-strategy", "XYZ"], "Valid strategy"),
                                          (["-type", "CPU"], "Not a single compatible device"),
                                          (["-cl-mls"], "needs a value")])
def test_acc_init_rejects_bad_options(tmp_path, argv, expect):
    """acc_interpret_init returns 1 with a message (accelerator.cc:565-606) before touching a device."""
    gr = ppiamd.Grammar(4)
    pop = ppiamd.grow(gr, 4, 1, 3, 16)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 1, 10, 4)
    inp = tmp_path / "in.bin"
    write_case(inp, data, pop)
    r = subprocess.run([os.path.join(SHIM, "abs", "harness"), "acc", str(inp), str(tmp_path / "out.bin")] + argv,
                       capture_output=True, text=True)
    assert r.returncode == 4, r.stderr
    assert expect in r.stderr


def test_grammar_table_compiled_into_shim():
    inc = open(os.path.join(SHIM, "abs", "inc", "ppi_symbol_table.inc")).read()
    assert inc.count("},") >= 14  # 6 binary + 6 unary + ternary + attribute + const + PI + 2 + N3
