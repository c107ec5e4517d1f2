"""C++ host layer: the input-deck reader (reference grammar, parser/Parser.y
and parser/Lexer.l) and the host program / header build."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("deck") / "deck_probe")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "a2c_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "deck_probe.cpp"), "-o", exe], check=True)
    return exe


def run(probe, tmp_path, text):
    p = tmp_path / "in.st"
    p.write_text(text)
    out = subprocess.run([probe, str(p)], capture_output=True, text=True).stdout
    if out.startswith("ERROR"):
        return out
    return {k: (kind, v) for k, kind, v in (line.split("\t") for line in out.splitlines())}


def test_deck_grammar(probe, tmp_path):
    text = """
    // comment
    under Diffusion {
      Temperature = 400;            /* int constant */
      AttemptFrequency = 1.0e13;
      ActivationEnergy = 0.23;
    }
    under Lattice[0] { PeriodicOA = True; Name = "pd_fcc"; }
    under Time { TimeStep = 5.0d-13; MaxIts = 2*50; }
    Output.Prefix = "results/";
    Mixed = (1 + 2) * 0.5;
    IntDiv = 7 / 2;
    Pow = 2^3^2;
    PowMul = 3 * 2^2;
    PowD = 2.0^0.5;
    PowNeg = 2^(0-1);
    end
    Ignored = 1;
    """
    v = run(probe, tmp_path, text)
    assert v["Diffusion.Temperature"] == ("int", "400")
    assert float(v["Diffusion.AttemptFrequency"][1]) == 1e13
    assert v["Lattice[0].PeriodicOA"] == ("sym", "True")
    assert v["Lattice[0].Name"] == ("str", "pd_fcc")
    assert float(v["Time.TimeStep"][1]) == 5e-13
    assert v["Time.MaxIts"] == ("int", "100")
    assert v["Output.Prefix"] == ("str", "results/")
    assert float(v["Mixed"][1]) == 1.5
    assert v["IntDiv"] == ("int", "3")  # IntExpr / IntExpr stays integer
    assert v["Pow"] == ("int", "64")  # %left '^': (2^3)^2
    assert v["PowMul"] == ("int", "12")  # '^' binds tighter than '*'
    assert float(v["PowD"][1]) == 2.0 ** 0.5
    assert float(v["PowNeg"][1]) == 0.5
    assert "Ignored" not in v


def test_deck_errors_name_the_line(probe, tmp_path):
    out = run(probe, tmp_path, "under A {\n B = ;\n}\n")
    assert isinstance(out, str) and "line 2" in out


def test_host_program_and_mirror_header_build():
    exe = os.path.join(ROOT, "a2c_amd", "a2c_dmd")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    # a caller written against the reference's signatures (minimizer.h:82-98)
    # compiles against the mirror header (it runs in the GPU tests)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "minimizer_caller.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
