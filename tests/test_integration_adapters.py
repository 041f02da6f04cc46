"""The reference-side adapters written in INTEGRATION.md compile (g++
-fsyntax-only) against include/sbs/sbs.h and declarations of the reference
interfaces they bind (tests/integration/stubs: the shapes of Solver,
LinearSolver, SolverData, SolverParams, FixedBodyData, glm vectors).  Every
code block whose first line names a reference file it creates or replaces is
checked; a block that stops compiling after a change of the C ABI fails
here."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STUBS = os.path.join(REPO, "tests", "integration", "stubs")
PRELUDE = """#include <cstring>
#include <memory>
#include <stdexcept>
#include <def.h>
#include <simulation/fixedBodyData.h>
"""


# explicit instantiations, so the templates' member bodies are checked too
EPILOGUE = {
    "SbsPCGJacobiSolver": "template class SbsPCGJacobiSolver<float>;\ntemplate class SbsPCGJacobiSolver<double>;\n",
    "SbsCollisionDetection": "template class SbsCollisionDetection<float>;\n"
                             "template class SbsCollisionDetection<double>;\n",
    "FixedBodyData::HandleCollisions": "template void FixedBodyData::HandleCollisions<float>(glm::tvec3<float>*, "
                                       "glm::tvec3<float>*, int, float, float);\n"
                                       "template void FixedBodyData::HandleCollisions<double>(glm::tvec3<double>*, "
                                       "glm::tvec3<double>*, int, double, double);\n",
}


def adapter_blocks():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    return [b for b in blocks if b.startswith("// src/")]


def test_integration_doc_has_the_adapters():
    heads = [b.splitlines()[0] for b in adapter_blocks()]
    for name in ("sbsPdSolver.h", "fixedBodyData.cu", "sbsPcgJacobi.h", "sbsCollisionDetection.h",
                 "sbsIpcSolver.h"):
        assert any(name in h for h in heads), (name, heads)


@pytest.mark.parametrize("idx", range(5))
def test_adapter_compiles(tmp_path, idx):
    blocks = adapter_blocks()
    assert idx < len(blocks)
    block = blocks[idx].replace("#pragma once\n", "")
    src = tmp_path / f"adapter{idx}.cpp"
    tail = "".join(v for k, v in EPILOGUE.items() if k in block)
    src.write_text(PRELUDE + block + tail)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(REPO, "include"),
                        "-I", STUBS, str(src)], capture_output=True, text=True)
    assert r.returncode == 0, blocks[idx].splitlines()[0] + "\n" + r.stderr[-3000:]
