"""molflowCLI for the GPU unit: the command line driver of the reference
(src/MolflowCLI.cpp:72-272, options of src/Initializer.cpp:31-100) running
the MI355X tracer.

    python -m molflow_sources_amd.cli -f model.xml -d 1e8 -o result.xml
    python -m molflow_sources_amd.cli -f result.xml -t 600 -a 60 -s 10

Loads a MolFlow XML (or zipped XML) file and, unless --reset, continues
from the results it holds (desorption limits count those too, and the new
particles take random streams after the ones already run). Runs until the
desorption limit (-d) or the time limit (-t) is reached, printing
statistics every -s seconds and writing an autosave file every -a seconds;
then writes the results file (MolflowResults in the reference's layout,
readable by MolFlow+) and, on request, the facet CSV exports. Parameter
changes (--setParams, --setParamsByFile) are applied before the run.

Differences from the reference, all deliberate: one GPU replaces the CPU
threads (-j is accepted and ignored); when a limit is reached the
particles in flight are traced to their end before the results are
written, so every desorbed particle is complete (the reference stops its
threads where they are on a time limit); MPI is not part of this build.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
import zipfile



def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="molflowCLI (MI355X)", description=__doc__.split("\n\n")[0])
    ap.add_argument("-f", "--file", required=True, help="input file (XML or ZIP)")
    ap.add_argument("-o", "--output", help="output file name, default out_<input file name>")
    ap.add_argument("--outputPath", help="output directory, default Results_<timestamp>")
    ap.add_argument("-d", "--ndes", type=float, default=0.0, help="finish after this many desorbed particles")
    ap.add_argument("-t", "--time", type=float, default=0.0, help="finish after this many seconds")
    ap.add_argument("-s", "--statprintInterval", type=float, default=60.0,
                    help="statistics print interval in seconds, 0 = disabled")
    ap.add_argument("-a", "--autosaveInterval", type=float, default=600.0,
                    help="autosave interval in seconds, 0 = disabled")
    ap.add_argument("-r", "--reset", action="store_true", help="start from 0 desorptions (drop stored results)")
    ap.add_argument("--setParams", action="append", default=[], help="parameter changes, e.g. facet.3.sticking=0.1")
    ap.add_argument("--setParamsByFile", help="file of parameter changes, one per line")
    ap.add_argument("--writeFacetDetails", action="store_true", help="write facet_details.csv")
    ap.add_argument("--writeFacetQuantities", action="store_true", help="write facet_physics.csv")
    ap.add_argument("--loadAutosave", action="store_true", help="continue from the autosave file if it exists")
    ap.add_argument("--noProgress", action="store_true", help="no progress lines")
    ap.add_argument("--overwrite", action="store_true", help="write the results into the input file")
    ap.add_argument("-j", "--threads", type=int, default=0, help="accepted for compatibility; the GPU runs every lane")
    ap.add_argument("--catalog", help="parameter catalog directory (CSV files)")
    ap.add_argument("--seed", type=int, default=42, help="random seed (Philox key)")
    ap.add_argument("--desPerStep", type=float, default=1e6, help="desorptions per GPU launch")
    ap.add_argument("--device", type=int, default=0)
    return ap


class StatPrinter:
    """RuntimeStatPrinter role: elapsed time, hits, desorptions and rates."""

    def __init__(self, hits0: int, des0: int):
        self.hits0, self.des0 = hits0, des0

    @staticmethod
    def header():
        print(f"{'Time':>12} {'#Hit':>16} {'#Des':>14} {'Hit/sec':>14} {'Des/sec':>14}", flush=True)

    def line(self, elapsed: float, hits: int, des: int):
        dt = max(elapsed, 1e-9)
        print(f"{elapsed:12.2f} {hits:16d} {des:14d} {(hits - self.hits0) / dt:14.4g} {(des - self.des0) / dt:14.4g}",
              flush=True)


def output_paths(args):
    base = os.path.basename(args.file)
    if args.overwrite:
        return os.path.dirname(os.path.abspath(args.file)) or ".", os.path.abspath(args.file)
    outdir = args.outputPath or f"Results_{time.strftime('%Y-%m-%d_%H-%M-%S')}"
    os.makedirs(outdir, exist_ok=True)
    name = args.output or ("out_" + base)
    return outdir, name if os.path.isabs(name) else os.path.join(outdir, name)


def write_results(path: str, doc, tallies, des_limit: int, totals) -> None:
    from . import physics, xml_io
    n_des = int(tallies.glob[0]["nb_desorbed"])
    mptp = (lambda m: physics.molecules_per_tp(m, totals, n_des, doc.model)) if totals else None
    if path.lower().endswith(".zip"):
        xml_path = path[:-4] + ".xml"
        xml_io.save(xml_path, doc, tallies, des_limit=des_limit, molecules_per_tp=mptp)
        with zipfile.ZipFile(path + ".tmp", "w", zipfile.ZIP_DEFLATED) as z:
            z.write(xml_path, os.path.basename(xml_path))
        os.replace(path + ".tmp", path)
        os.remove(xml_path)
    else:
        xml_io.save(path, doc, tallies, des_limit=des_limit, molecules_per_tp=mptp)


def main(argv=None) -> int:
    """Exit codes as the reference's: 41 bad arguments or no stop
    condition, 42 bad parameter changes or model, 43 file that does not
    load or run that fails, 0 success."""
    try:
        args = build_parser().parse_args(argv)
    except SystemExit as e:
        return 0 if e.code == 0 else 41
    from . import params as P
    from . import physics, xml_io
    from .simulation import GpuSimulation, MolflowGpuError, sanity_check
    from .tallies import Tallies

    if args.overwrite and (args.output or args.outputPath):
        print("--overwrite excludes --output and --outputPath", file=sys.stderr)
        return 41
    if args.ndes < 0:
        print(f"Invalid desorption limit {args.ndes:.4g}, must be larger than 0", file=sys.stderr)
        return 41
    if args.ndes <= 0 and args.time <= 0:
        print("No stop condition set. Use either -t (--time) or -d (--ndes)", file=sys.stderr)
        return 41
    if not os.path.isfile(args.file):
        print(f"--file: File does not exist: {args.file}", file=sys.stderr)
        return 41
    outdir, out_file = output_paths(args)
    autosave = os.path.join(outdir, "autosave_" + os.path.basename(args.file).rsplit(".", 1)[0] + ".xml")
    src = autosave if args.loadAutosave and os.path.exists(autosave) else args.file
    try:
        doc = xml_io.load(src, catalog_dir=args.catalog, with_results=not args.reset)
    except Exception as e:                         # noqa: BLE001 -- reported like the loader's errors
        print(f"Error loading file:\n{e}", file=sys.stderr)
        return 43
    model = doc.model
    changes = P.parse(args.setParams, P.selections_of(doc.root))
    if args.setParamsByFile:
        more = P.parse_file(args.setParamsByFile, P.selections_of(doc.root))
        changes.facet += more.facet
        changes.simulation += more.simulation
        changes.messages += more.messages
    if changes.facet or changes.simulation or changes.messages:
        if P.apply(model, changes):
            print("Invalid parameter changes", file=sys.stderr)
            return 42
        doc.results = None if args.reset else doc.results
    problems = sanity_check(model)
    if problems:
        print("Model check failed:\n" + "\n".join(problems), file=sys.stderr)
        return 42
    totals = physics.model_totals(model)
    sim = GpuSimulation(args.device)
    sim.load(model)
    sim.rebuild_accel()
    prior = doc.results
    des0 = int(prior.glob[0]["nb_desorbed"]) if prior is not None else 0
    hits0 = int(prior.glob[0]["nb_mc_hit"]) if prior is not None else 0
    # new particles take the random streams after the ones already run
    sim.reset(seed=args.seed, particle_offset=des0)
    sim.set_carry(True)
    limit = int(args.ndes) if args.ndes > 0 else 0
    if limit and des0 >= limit:
        print(f"Desorption limit ({limit}) already reached (input file nbDes={des0}). "
              "Consider resetting the simulation (--reset argument)", flush=True)
    per = max(1, int(args.desPerStep))
    if args.time > 0:
        sim.set_time_limit(max(1.0, min(60.0, args.time)))
    acc = Tallies(sim.layout(), sim.offsets())
    if prior is not None:
        acc += prior
    printer = StatPrinter(hits0, des0)
    if limit:
        print(f"Starting simulation to {limit} desorptions from {des0} desorptions...", flush=True)
    else:
        print(f"Starting simulation for {args.time} seconds from {des0} desorptions...", flush=True)
    t0 = time.perf_counter()
    last_save = last_stat = t0
    if args.statprintInterval and not args.noProgress:
        printer.header()
    while True:
        done = int(acc.glob[0]["nb_desorbed"])
        now = time.perf_counter()
        if (limit and done >= limit) or (args.time > 0 and now - t0 >= args.time):
            break
        n = per if not limit else min(per, limit - done)
        try:
            sim.run(n)
        except MolflowGpuError as e:
            print(f"ERROR: {e}", file=sys.stderr)
            return 43
        sim.take(acc)                              # periodic report: add into the host state, clear the device
        now = time.perf_counter()
        if args.statprintInterval and not args.noProgress and now - last_stat >= args.statprintInterval:
            printer.line(now - t0, int(acc.glob[0]["nb_mc_hit"]), int(acc.glob[0]["nb_desorbed"]))
            last_stat = now
        if args.autosaveInterval and now - last_save >= args.autosaveInterval:
            write_results(autosave, doc, acc, limit, totals)
            last_save = now
    # complete the particles still in flight, then the last report
    sim.set_time_limit(60.0)
    sim.drain()
    sim.take(acc)
    elapsed = time.perf_counter() - t0
    sim.close()
    print("Simulation finished.", flush=True)
    printer.header()
    printer.line(elapsed, int(acc.glob[0]["nb_mc_hit"]), int(acc.glob[0]["nb_desorbed"]))
    if args.writeFacetDetails:
        physics.write_facet_details_csv(os.path.join(outdir, "facet_details.csv"), model, acc, totals)
    if args.writeFacetQuantities:
        physics.write_facet_quantities_csv(os.path.join(outdir, "facet_physics.csv"), model, acc, totals)
    write_results(out_file, doc, acc, limit, totals)
    if os.path.exists(autosave) and src != autosave:
        os.remove(autosave)
    print(f"Results written to {out_file}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
