"""Sweep writers (host code, no GPU): the MAT v4 bytes of Zc.mat
(fasterhenry-cli/src/mat.rs), the SPICE subcircuit stamp
(fasterhenry-cli/src/spice.rs) and the subcircuit name of an input path,
against the layout and cases of the reference's own tests
(fasterhenry-cli/tests/roundtrip.rs:73-142, spice.rs:128-160) and SciPy's
MAT reader."""
import io
import struct

import numpy as np
import scipy.io

from fasterhenry_amd import spice_subckt, subckt_name_from_path, zc_mat


def _sweep(K, p, seed=0):
    rng = np.random.default_rng(seed)
    f = np.concatenate([[0.0], np.logspace(6, 9, K - 1)]) if K > 1 else np.array([1e6])
    z = rng.standard_normal((K, p, p)) + 1j * rng.standard_normal((K, p, p))
    z = 0.5 * (z + z.transpose(0, 2, 1))
    z[0] = z[0].real if f[0] == 0.0 else z[0]
    return f, z


def test_mat4_layout_of_a_two_frequency_one_port_sweep():
    f, z = _sweep(2, 1)
    b = zc_mat(f, z)
    # freqs (1 x 2 real): 20 + 5 + 16; Zc_1, Zc_2 (1 x 1 complex): 20 + 4 + 16 each
    assert len(b) == 41 + 40 + 40
    assert b[0:4] == struct.pack("<i", 0) and b[41:45] == struct.pack("<i", 0)
    assert b[81:85] == struct.pack("<i", 0)
    assert b[12:16] == struct.pack("<i", 0)  # imagf of freqs
    assert b[53:57] == struct.pack("<i", 1)  # imagf of Zc_1
    assert b[20:25] == b"freqs"
    assert struct.unpack("<2d", b[25:41]) == tuple(f)
    assert b[61:65] == b"Zc_1"
    assert struct.unpack("<2d", b[65:81]) == (z[0, 0, 0].real, z[0, 0, 0].imag)


def test_mat4_reads_back_with_scipy():
    f, z = _sweep(5, 3, seed=2)
    m = scipy.io.loadmat(io.BytesIO(zc_mat(f, z)))
    assert np.array_equal(m["freqs"], f[None, :])
    for k in range(5):
        assert np.array_equal(m[f"Zc_{k + 1}"], z[k])  # column-major, real then imaginary


def test_spice_subckt_stamps_r_and_l():
    f = np.array([1e6])
    z = np.array([[[0.25 + 2.0j]]])
    text = spice_subckt(f, z, 0, "spiral")
    assert ".subckt spiral p1 n1\n" in text and text.endswith(".ends spiral\n")
    assert "L1 p1 a0_0 " in text and "vsense1 a0_1 n1 0" in text
    assert "H1_1 a0_0 a0_1 vsense1 " in text and "\nK" not in text
    assert "* fasterhenry SPICE export: spiral at 1.00000000000000000e6 Hz\n" in text
    lval = float(text.split("L1 p1 a0_0 ")[1].split("\n")[0])
    assert lval == 2.0 / (6.283185307179586 * 1e6)
    # Rust's {:.17e}: no '+', no zero-padded exponent
    assert "H1_1 a0_0 a0_1 vsense1 2.50000000000000000e-1\n" in text


def test_spice_two_ports_carry_k_and_off_diagonal_h():
    f = np.array([0.0, 1e6])
    w = 6.283185307179586e6
    z = np.zeros((2, 2, 2), dtype=complex)
    z[0] = [[0.1, 0.02], [0.02, 0.1]]
    z[1] = [[0.1 + 1j * w * 2e-9, 0.02 + 1j * w * 5e-10], [0.02 + 1j * w * 5e-10, 0.1 + 1j * w * 2e-9]]
    text = spice_subckt(f, z, 1, "twobar")
    assert "K1_2 L1 L2 " in text
    k = float(text.split("K1_2 L1 L2 ")[1].split("\n")[0])
    assert abs(k - 0.25) < 1e-15
    assert "H1_2 a0_1 a0_2 vsense2 " in text and "H2_1 a1_0 a1_1 vsense1 " in text
    # DC: inductors are wires, no K lines
    dc = spice_subckt(f, z, 0, "twobar")
    assert "L1 p1 a0_0 0.00000000000000000e0\n" in dc and "\nK" not in dc
    # a coupling a hair above 1 is clamped to 1 - 1e-12
    z[1, 0, 1] = z[1, 1, 0] = 0.02 + 1j * w * 2.0000000001e-9
    t = spice_subckt(f, z, 1, "twobar")
    assert float(t.split("K1_2 L1 L2 ")[1].split("\n")[0]) == 1.0 - 1e-12


def test_subckt_names_follow_the_reference_cases():
    # fasterhenry-cli/src/spice.rs:128-160
    assert subckt_name_from_path("spiral.inp") == "spiral"
    assert subckt_name_from_path("dir/two_turn_2.json") == "two_turn_2"
    assert subckt_name_from_path("my coil.inp") == "my_coil"
    assert subckt_name_from_path("a\tb\nc.inp") == "a_b_c"
    assert subckt_name_from_path("v1.2-final(x).inp") == "v1_2_final_x_"
    assert subckt_name_from_path("bobineé.inp") == "bobine_"
    assert subckt_name_from_path("線圈.inp") == "__"
    assert subckt_name_from_path("x\n.ends\n.subckt evil.inp") == "x__ends__subckt_evil"
    assert subckt_name_from_path("") == "fasterhenry"
    assert subckt_name_from_path("/") == "fasterhenry"
    assert subckt_name_from_path("..") == "fasterhenry"
