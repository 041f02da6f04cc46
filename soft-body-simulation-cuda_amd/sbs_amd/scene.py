"""Headless scene configuration: the reference's JSON scene file (context.json)
loaded into solvers of this library, with per-phase timers.

Mirrors, without the OpenGL/ImGui parts:
  * Context::LoadSimContext (src/context.cpp:318-380): the top-level keys
    "num of iterations" (default 10) and "threads per block" (default 128), the
    "softBodies" and "fixedBodies" definitions by name, and one simulation
    context per "contexts" entry whose "load" is not false;
  * SimulationCUDAContext::Impl::Init (src/simulation/simulationContext.cu:35-123):
    the context's "dt", "gravity", "tolerance", "maxIterations", "dhat",
    "damp", "muN", "muT", "kappa" and "precision" ("float"/"fp32": the PD
    solver; anything else: the FP64 IPC solver); each soft body instance takes
    "pos", "scale", "rot" from itself, else from its definition, else 0
    (getJsonVec3, :24-31, also for scale); "mass", "mu", "lambda" likewise with
    defaults 1, 1000, 1000; "DBC" from the instance, else the definition;
    "centralize", "start index", "nodeFile", "eleFile", "faceFile" from the
    definition ("mshFile" bodies are not supported: no Gmsh reader here);
  * ReadFixedBodies (src/context.cpp:222-316): sphere / cylinder / plane with
    pos, rot, scale (defaults 0, 0, 1) and the sphere radius;
  * SimulationCUDAContext::Update / Reset / GetPerformanceData
    (simulationContext.cpp:80-95, simulationContext.cu:215-224,
    simulationContext.cpp:160-172), with the float tolerance clamped at 1e-6
    (CopyUIToParams, simulationContext.cpp:29-32) and collision handling on
    (the GUI's default, context.h:42-44).

Relative file names are resolved against `base_dir` (the directory of the JSON
file by default), then by file name in each of `search_dirs`.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import fixed as F
from .loader import DataLoader, SoftBodyAttribute
from .pd import PdSolver, SolverData, SolverParams, SolverType


def _vec3(inst: dict, defn: dict, key: str, default=(0.0, 0.0, 0.0)):
    src = inst if key in inst else defn if key in defn else None
    return tuple(float(v) for v in src[key]) if src is not None else tuple(default)


def _val(inst: dict, defn: dict, key: str, default):
    return inst[key] if key in inst else defn.get(key, default)


class SceneFile:
    """A parsed scene file: definitions and the loadable contexts."""

    def __init__(self, path: str, base_dir: str | None = None, search_dirs: tuple = ()):
        self.path = path
        with open(path) as f:
            self.json = json.load(f)
        self.base_dir = base_dir if base_dir is not None else os.path.dirname(os.path.abspath(path))
        self.search_dirs = tuple(search_dirs)
        j = self.json
        self.num_iterations = int(j.get("num of iterations", 10))
        self.threads_per_block = int(j.get("threads per block", 128))
        self.soft_defs = {d["name"]: d for d in j.get("softBodies", [])}
        self.fixed_defs = {d["name"]: d for d in j.get("fixedBodies", [])}
        self.contexts = [c for c in j.get("contexts", []) if c.get("load", True)]

    def resolve(self, name: str) -> str:
        if not name:
            return ""
        cands = [name] if os.path.isabs(name) else [os.path.join(self.base_dir, name)]
        cands += [os.path.join(d, os.path.basename(name)) for d in self.search_dirs]
        for c in cands:
            if os.path.exists(c):
                return c
        raise FileNotFoundError(f"{name}: not found (looked in {cands})")

    def context(self, name: str | int = 0) -> dict:
        if isinstance(name, int):
            return self.contexts[name]
        for c in self.contexts:
            if c["name"] == name:
                return c
        raise KeyError(f"no loadable context named {name!r}")

    def fixed_bodies(self, ctx: dict) -> list:
        out = []
        for inst in ctx.get("fixedBodies", []):
            d = self.fixed_defs[inst["name"]]
            pos = _vec3(inst, d, "pos")
            rot = _vec3(inst, d, "rot")
            scale = _vec3(inst, d, "scale", (1.0, 1.0, 1.0))
            kind = d["type"]
            if kind == "sphere":
                r = float(_val(inst, d, "radius", 1.0))
                out.append(F.Sphere(pos, rot, r, inst["name"]))
            elif kind == "cylinder":
                out.append(F.Cylinder(pos, rot, scale, inst["name"]))
            elif kind == "plane":
                out.append(F.Plane(pos, rot, scale, inst["name"]))
            else:
                raise ValueError(f"fixed body {inst['name']!r}: unknown type {kind!r}")
        return out

    def soft_scene(self, ctx: dict, precision: int):
        dl = DataLoader(self.threads_per_block, precision)
        for inst in ctx.get("softBodies", []):
            d = self.soft_defs[inst["name"]]
            if d.get("mshFile"):
                raise NotImplementedError(f"{inst['name']}: .msh bodies are not supported")
            if not d.get("nodeFile"):
                raise ValueError(f"{inst['name']}: a node file must be provided")
            attr = SoftBodyAttribute(mass=float(_val(inst, d, "mass", 1.0)), mu=float(_val(inst, d, "mu", 1000.0)),
                                     lam=float(_val(inst, d, "lambda", 1000.0)),
                                     DBC=[int(v) for v in (inst["DBC"] if "DBC" in inst else d.get("DBC") or [])])
            dl.CollectData(self.resolve(d["nodeFile"]), self.resolve(d.get("eleFile", "")),
                           self.resolve(d["faceFile"]) if d.get("faceFile") else None,
                           _vec3(inst, d, "pos"), _vec3(inst, d, "scale"), _vec3(inst, d, "rot"),
                           bool(d.get("centralize", False)), int(d.get("start index", 0)), attr, inst["name"])
        return dl.AllocData(), dl


class SimulationContext:
    """One simulation context of a scene file on the GPU (PD for float
    contexts, IPC for double ones)."""

    def __init__(self, scene_file: SceneFile, name: str | int = 0, global_solver: SolverType = SolverType.Jacobi,
                 handle_collision: bool = True, num_iterations: int | None = None):
        sf = scene_file
        ctx = sf.context(name)
        self.name = ctx["name"]
        prec = ctx.get("precision", "double")
        self.precision = 32 if prec in ("float", "fp32") else 64
        self.scene, self.loader = sf.soft_scene(ctx, self.precision)
        self.fixed_specs = sf.fixed_bodies(ctx)
        if self.precision == 32:
            tol = max(float(ctx.get("tolerance", 1e-2)), 1e-6)
            self.params = SolverParams(numIterations=num_iterations or sf.num_iterations,
                                       dt=float(ctx.get("dt", 0.001)), gravity=float(ctx.get("gravity", 9.8)),
                                       tol=tol, handleCollision=handle_collision,
                                       muN=float(ctx.get("muN", 0.5)), muT=float(ctx.get("muT", 0.5)))
            self.data = SolverData(self.scene, 32)
            self.fixed = F.FixedBodyData(self.fixed_specs) if self.fixed_specs else None
            self.data.set_fixed_bodies(self.fixed)
            self.solver = PdSolver(sf.threads_per_block, self.data)
            self.solver.SetGlobalSolver(global_solver)
            self.solver.SetPerf(True)
        else:
            from .ipc import IpcData, IpcParams, IpcSolver
            # fixed bodies enter the IPC energy as its implicit barrier
            self.params = IpcParams(dt=float(ctx.get("dt", 0.001)), damp=float(ctx.get("damp", 0.999)),
                                    dhat=float(ctx.get("dhat", 1e-2)), tol=float(ctx.get("tolerance", 1e-2)),
                                    maxIterations=int(ctx.get("maxIterations", 100)))
            self.fixed = F.FixedBodyData(self.fixed_specs) if self.fixed_specs else None
            self.data = IpcData(self.scene, fixed=self.fixed)
            self.solver = IpcSolver(self.data)
        self.steps = 0
        self.wall_ms = 0.0

    def Update(self) -> bool:
        t0 = time.perf_counter()
        ok = self.solver.Update(self.data, self.params)
        if self.precision == 32:
            self.solver.synchronize()
        self.wall_ms += (time.perf_counter() - t0) * 1e3
        self.steps += 1
        return ok if ok is not None else True

    def Reset(self) -> None:
        self.solver.Reset()
        self.steps = 0
        self.wall_ms = 0.0

    def GetPerformanceData(self) -> list[tuple[str, float]]:
        if self.precision == 32:
            return self.solver.GetPerformanceData()
        st = self.solver.stats()
        return [("newton iterations", float(st["newton_iterations"]))]

    def positions(self) -> np.ndarray:
        return self.data.get("X")

    def close(self) -> None:
        self.solver.close()
        if self.fixed is not None:
            self.fixed.close()
        self.data.free()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="run a scene file's context headless and print per-phase timers")
    ap.add_argument("scene")
    ap.add_argument("--context", default="0", help="context name or index")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--assets", action="append", default=[], help="extra directories to find mesh files in")
    ap.add_argument("--solver", default="Jacobi", choices=[s.name for s in SolverType])
    ap.add_argument("--iterations", type=int, default=None, help="override 'num of iterations'")
    a = ap.parse_args(argv)
    sf = SceneFile(a.scene, search_dirs=tuple(a.assets))
    name = int(a.context) if a.context.isdigit() else a.context
    sim = SimulationContext(sf, name, SolverType[a.solver], num_iterations=a.iterations)
    try:
        for _ in range(a.steps):
            sim.Update()
        print(json.dumps({"context": sim.name, "precision": sim.precision, "verts": sim.scene.nv,
                          "tets": sim.scene.nt, "tris": sim.scene.ntri, "fixed_bodies": len(sim.fixed_specs),
                          "steps": sim.steps, "ms_per_step_wall": sim.wall_ms / max(sim.steps, 1),
                          "phases_ms": dict(sim.GetPerformanceData())}))
    finally:
        sim.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
