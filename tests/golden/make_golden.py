"""Extracts the reference's own unit-test tables for the FV RHS leaf functions into JSON fixtures.

Run in a container that has the ABLATE source tree (default /root/reference); the fixtures it writes are data
(inputs and expected outputs), committed next to this script:
  flux_calculators.json   fluxCalculatorTests.cpp: Ausm, AverageFlux and OffFlux tables
  ausmpup.json            fluxCalculatorTests.cpp: AusmpUp tables with their constructor arguments (mInf, and the
                          initial alpha of the PressureGradientScaling object when one is attached)
  ns_advection_flux.json  navierStokesTransportTests.cpp: AdvectionFlux cases
  stress_tensor.json      navierStokesTransportTests.cpp: CompressibleFlowComputeStressTensor cases
  perfect_gas.json        perfectGasTests.cpp: PerfectGas thermodynamic property cases
  shock_tube.json         compressibleShockTubeTests.cpp: first-order AUSM shock-tube cases (forward Euler, CFL
                          time step) with the reference's x, rho, u, e profiles
  pipeFlow.msh            tests/integrationTests/inputs/compressibleFlow/gmshPipeFlow/pipeFlow.msh (data, copied)
  mixedCells2D.msh        tests/integrationTests/inputs/compressibleFlow/mixedCellTypeTest/mixedCells2D.msh (copied)
  2D_Circle.hdf5, 2D_Circle_Tri.hdf5
                          tests/integrationTests/inputs/levelSet/ (Hdf5Serializer output of an 80 x 80 box on
                          [-2, 2]^2, quadrilaterals / triangles, with the volume fraction of the unit circle; copied)
  initializer.2D.hdf5, initializer.3D.hdf5
                          tests/unitTests/inputs/domain/ (copied)
  hdf5_initializer.json   hdf5InitializerTests.cpp: the points and expected field values of its test table
Each record keeps the source file and line of its table so that a reader can find it.
"""
from __future__ import annotations

import json
import math
import re
import shutil
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent
UT = REF / "tests" / "unitTests"

_NUM = re.compile(r"^[\s0-9eE+\-.*/()]*$")


def num(expr: str) -> float:
    """Evaluates a numeric C++ literal expression of the tables (literals, * and /, NAN)."""
    expr = expr.strip()
    if expr in ("NAN", "(double)NAN"):
        return math.nan
    if not _NUM.match(expr):
        raise ValueError(f"unexpected expression {expr!r}")
    return float(eval(expr, {"__builtins__": {}}, {}))  # arithmetic literal only (checked above)


def braces(text: str, start: int) -> tuple[str, int]:
    """Returns the content of the {...} block opening at text[start] and the index after it."""
    assert text[start] == "{"
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start + 1 : i], i + 1
    raise ValueError("unbalanced braces")


def field_list(block: str, name: str):
    m = re.search(r"\." + name + r"\s*=\s*(std::vector<PetscReal>)?\s*\{", block)
    if not m:
        return None
    inner, _ = braces(block, m.end() - 1)
    items = [x for x in inner.replace("\n", " ").split(",") if x.strip()]
    return items


def line_of(text: str, idx: int) -> int:
    return text.count("\n", 0, idx) + 1


def flux_calculators():
    path = UT / "finiteVolume" / "fluxCalculator" / "fluxCalculatorTests.cpp"
    text = path.read_text()
    out = {}
    for m in re.finditer(r"\(FluxCalculatorTestParameters\)\s*\{", text):
        block, _ = braces(text, m.end() - 1)
        name = re.search(r'\.testName\s*=\s*"([^"]+)"', block).group(1)
        if name not in ("Ausm", "AverageFlux", "OffFlux"):
            continue
        rec = {"source": f"tests/unitTests/finiteVolume/fluxCalculator/fluxCalculatorTests.cpp:{line_of(text, m.start())}"}
        for key in ("uL", "aL", "rhoL", "pL", "uR", "aR", "rhoR", "pR", "expectedMassFlux", "expectedInterfacePressure"):
            rec[key] = [num(x) for x in field_list(block, key)]
        rec["expectedDirection"] = [x.strip() for x in field_list(block, "expectedDirection")]
        rec["tolerance"] = {"massFlux": 1e-2, "interfacePressure": 1.0}
        out[name] = rec
    return out


def ausmpup_tables():
    path = UT / "finiteVolume" / "fluxCalculator" / "fluxCalculatorTests.cpp"
    text = path.read_text()
    out = {}
    for m in re.finditer(r"\(FluxCalculatorTestParameters\)\s*\{", text):
        block, _ = braces(text, m.end() - 1)
        name = re.search(r'\.testName\s*=\s*"([^"]+)"', block).group(1)
        if not name.startswith("AusmpUp"):
            continue
        ctor = re.search(r"AusmpUp>\(\s*([^,)]+)", block)
        pgs = re.search(r"PressureGradientScaling>\(\s*[^,]+,\s*([^,]+),", block)
        rec = {"source": f"tests/unitTests/finiteVolume/fluxCalculator/fluxCalculatorTests.cpp:{line_of(text, m.start())}",
               "mInf": num(ctor.group(1)), "pgsAlpha": num(pgs.group(1)) if pgs else None}
        for key in ("uL", "aL", "rhoL", "pL", "uR", "aR", "rhoR", "pR", "expectedMassFlux", "expectedInterfacePressure"):
            rec[key] = [num(x) for x in field_list(block, key)]
        rec["expectedDirection"] = [x.strip() for x in field_list(block, "expectedDirection")]
        rec["tolerance"] = {"massFlux": 1e-2, "interfacePressure": 1.0}
        out[name] = rec
    return out


def ns_cases():
    path = UT / "finiteVolume" / "processes" / "navierStokesTransportTests.cpp"
    text = path.read_text()
    adv, stress = [], []
    for m in re.finditer(r"\(NavierStokesTransportFluxTestParameters\)\s*\{", text):
        block, _ = braces(text, m.end() - 1)
        flux = re.search(r"fluxCalculator::(\w+)>", block).group(1)
        adv.append({
            "source": f"tests/unitTests/finiteVolume/processes/navierStokesTransportTests.cpp:{line_of(text, m.start())}",
            "fluxCalculator": flux,
            "area": [num(x) for x in field_list(block, "area")],
            "xLeft": [num(x) for x in field_list(block, "xLeft")],
            "xRight": [num(x) for x in field_list(block, "xRight")],
            "expectedFlux": [num(x) for x in field_list(block, "expectedFlux")],
            "tolerance": 1e-3,
            "eos": {"gamma": 1.4, "Rgas": 287.0},
        })
    for m in re.finditer(r"\(StressTensorTestParameters\)\s*\{", text):
        block, _ = braces(text, m.end() - 1)
        stress.append({
            "source": f"tests/unitTests/finiteVolume/processes/navierStokesTransportTests.cpp:{line_of(text, m.start())}",
            "dim": int(re.search(r"\.dim\s*=\s*(\d+)", block).group(1)),
            "mu": num(re.search(r"\.mu\s*=\s*([0-9.eE+\-]+)", block).group(1)),
            "gradVel": [num(x) for x in field_list(block, "gradVel")],
            "expectedStressTensor": [num(x) for x in field_list(block, "expectedStressTensor")],
            "tolerance": 1e-8,
        })
    return adv, stress


def perfect_gas():
    path = UT / "eos" / "perfectGasTests.cpp"
    text = path.read_text()
    out = []
    for m in re.finditer(r"\(PGTestParameters\)\s*\{", text):
        block, _ = braces(text, m.end() - 1)
        opts = dict(re.findall(r'\{"(\w+)",\s*"([^"]+)"\}', block))
        prop = re.search(r"ThermodynamicProperty::(\w+)", block).group(1)
        fm = re.search(r'MockField::Create\("euler",\s*(\d+)(?:,\s*(\d+))?\)', block)
        species = re.search(r"\.species\s*=\s*(std::vector<std::string>)?\{([^}]*)\}", block)
        sp = [s.strip().strip('"') for s in species.group(2).split(",") if s.strip()] if species else []
        tm = re.search(r"\.expectedTemperature\s*=\s*([^,}\n]+)", block)
        T = None
        if tm and tm.group(1).strip() not in ("{}", "{"):
            T = num(tm.group(1))
        out.append({
            "source": f"tests/unitTests/eos/perfectGasTests.cpp:{line_of(text, m.start())}",
            "gamma": float(opts.get("gamma", 1.4)),
            "Rgas": float(opts.get("Rgas", 287.0)),
            "species": sp,
            "property": prop,
            "eulerComponents": int(fm.group(1)),
            "eulerOffset": int(fm.group(2) or 0),
            "conservedValues": [num(x) for x in field_list(block, "conservedValues")],
            "temperature": T,
            "expectedValue": [num(x) for x in field_list(block, "expectedValue")],
            "tolerance": 1e-6,
        })
    return out


def shock_tube():
    path = UT / "finiteVolume" / "compressibleShockTubeTests.cpp"
    text = path.read_text()
    out = []
    for m in re.finditer(r"\(CompressibleShockTubeParameters\)\s*\{", text):
        block, _ = braces(text, m.end() - 1)
        ic_m = re.search(r"\.initialConditions\s*=\s*\{", block)
        ic_block, _ = braces(block, ic_m.end() - 1)
        ic = {k: num(v) for k, v in re.findall(r"\.(\w+)\s*=\s*([^,}]+)", ic_block)}
        ev_m = re.search(r"\.expectedValues\s*=\s*\{", block)
        ev_block, _ = braces(block, ev_m.end() - 1)
        expected = {}
        for vm in re.finditer(r'\{"(\w+)",\s*\{', ev_block):
            inner, _ = braces(ev_block, vm.end() - 1)
            expected[vm.group(1)] = [num(x) for x in inner.replace("\n", " ").split(",") if x.strip()]
        out.append({
            "source": f"tests/unitTests/finiteVolume/compressibleShockTubeTests.cpp:{line_of(text, m.start())}",
            "name": re.search(r'MpiTestParameter\("([^"]+)"\)', block).group(1),
            "fluxCalculator": re.search(r"fluxCalculator::(\w+)>", block).group(1),
            "initialConditions": ic,
            "nx": int(re.search(r"\.nx\s*=\s*(\d+)", block).group(1)),
            "maxTime": num(re.search(r"\.maxTime\s*=\s*([0-9.eE+\-]+)", block).group(1)),
            "cfl": num(re.search(r"\.cfl\s*=\s*([0-9.eE+\-]+)", block).group(1)),
            "expectedValues": expected,
            "tolerance": 1e-6,
            "timeStepper": {"ts_adapt_type": "physics", "ts_adapt_safety": 1.0, "ts_exact_final_time": "matchstep"},
        })
    return out


def hdf5_initializer_points():
    """hdf5InitializerTests.cpp INSTANTIATE_TEST_SUITE_P table: file -> field -> [(point, expected values)]."""
    src = UT / "domain" / "hdf5InitializerTests.cpp"
    text = src.read_text()
    start = text.index("INSTANTIATE_TEST_SUITE_P")
    line0 = text[:start].count("\n") + 1
    out = {"source": f"tests/unitTests/domain/hdf5InitializerTests.cpp:{line0}", "files": {}}
    chunks = text[start:].split('.hdf5File = "')[1:]
    for ch in chunks:
        fname = ch[: ch.index('"')].split("/")[-1]
        fields = {}
        parts = re.split(r'\{"(\w+)",', ch)
        for name, body in zip(parts[1::2], parts[2::2]):
            pts = re.findall(r"\.point = \{([^}]*)\}, \.expectedValue = \{([^}]*)\}", body)
            fields[name] = [{"point": [num(x) for x in pnt.split(",")], "expected": [num(x) for x in ev.split(",")]}
                            for pnt, ev in pts]
        out["files"][fname] = fields
    return out


def main():
    st = shock_tube()
    (OUT / "shock_tube.json").write_text(json.dumps(st, indent=1))
    fc = flux_calculators()
    adv, stress = ns_cases()
    pgt = perfect_gas()
    (OUT / "flux_calculators.json").write_text(json.dumps(fc, indent=1))
    (OUT / "ausmpup.json").write_text(json.dumps(ausmpup_tables(), indent=1))
    (OUT / "ns_advection_flux.json").write_text(json.dumps(adv, indent=1))
    (OUT / "stress_tensor.json").write_text(json.dumps(stress, indent=1))
    (OUT / "perfect_gas.json").write_text(json.dumps(pgt, indent=1, allow_nan=True))
    for rel in ("compressibleFlow/gmshPipeFlow/pipeFlow.msh", "compressibleFlow/mixedCellTypeTest/mixedCells2D.msh"):
        shutil.copyfile(REF / "tests" / "integrationTests" / "inputs" / rel, OUT / Path(rel).name)
    for rel in ("integrationTests/inputs/levelSet/2D_Circle.hdf5", "integrationTests/inputs/levelSet/2D_Circle_Tri.hdf5",
                "unitTests/inputs/domain/initializer.2D.hdf5", "unitTests/inputs/domain/initializer.3D.hdf5"):
        shutil.copyfile(REF / "tests" / rel, OUT / Path(rel).name)
    (OUT / "hdf5_initializer.json").write_text(json.dumps(hdf5_initializer_points(), indent=1))
    print({k: len(v["uL"]) for k, v in fc.items()}, len(adv), len(stress), len(pgt), len(st))


if __name__ == "__main__":
    main()
