"""Workloads for kernel profiles, and summaries of rocprofv3 output.

Run under rocprofv3 (kernel trace or counters), for example

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o trace -- \
        python3 -m biblbm_amd.profile_kernels collide rbc_in_pipe f64 40
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT_F -o pmc -- \
        python3 -m biblbm_amd.profile_kernels collide rbc_in_pipe f64 10 copy 268435456 5

and summarise with

    python3 -m biblbm_amd.profile_kernels stats OUT
    python3 -m biblbm_amd.profile_kernels traffic OUT_F OUT_W 268435456 k_collide_stream

Launch shapes come from the saved tuning table ($BIBLBM_TUNE_CACHE) when it
has the lattice's size, so that a bench run followed by a profile with the
same table profiles the shapes the bench timed.

`collide W S N` builds workload W (rbc_in_pipe, suspension,
suspension_interactions, taylor_green or capsule_in_shear) with storage S, runs two coupled steps so that the force
tiles are those of a running simulation, then N collide-and-stream launches.
`coupled W S N` runs N eager coupled steps instead; `ensemble B N [EO OO EN ON]`
steps the bench's ensemble of B shear boxes N times, batched, with the
given collide orders and nontemporal levels; `collisions W S N` runs
N collision-handler passes (Morse r_min 1, beta 1, strength 1e-3) on the
vertices of workload W after two coupled steps.  `copy B N` runs N
launches of a plain 16-byte-per-thread device copy of B bytes (the
calibration kernel for the HBM byte counters).
"""
import csv
import glob
import os
import sys


def _workload(name, storage):
    from . import workloads
    return getattr(workloads, name)(storage)


def run(argv):
    from . import capi, tuning
    capi.call("bl_init", 0)
    # lattices take the tuned launch shapes of their size from the saved
    # tuning table, when there is one ($BIBLBM_TUNE_CACHE)
    tuning.load()
    i = 0
    while i < len(argv):
        cmd = argv[i]
        if cmd in ("collide", "coupled"):
            W = _workload(argv[i + 1], argv[i + 2])
            n = int(argv[i + 3])
            W.ibm.coupled_steps(2, use_graph=False)
            capi.call("bl_sync")
            if cmd == "collide":
                ms = W.lattice.time_collide(n)
            else:
                import time
                t0 = time.perf_counter()
                W.ibm.coupled_steps(n, use_graph=False)
                capi.call("bl_sync")
                ms = (time.perf_counter() - t0) * 1e3
            print("%s %s %s: %d launches, %.4f ms each" % (cmd, argv[i + 1], argv[i + 2], n, ms / n), flush=True)
            W.ibm.close()
            W.lattice.close()
            i += 4
        elif cmd == "collisions":
            import time
            from .potentials import MorsePotential
            W = _workload(argv[i + 1], argv[i + 2])
            n = int(argv[i + 3])
            u = MorsePotential(1.0, 1.0, 1e-3)
            W.ibm.coupled_steps(2, use_graph=False)
            capi.call("bl_sync")
            t0 = time.perf_counter()
            for _ in range(n):
                W.ibm.compute_collision_forces(u)
            capi.call("bl_sync")
            ms = (time.perf_counter() - t0) * 1e3
            print("collisions %s %s: %d passes, %.4f ms each" % (argv[i + 1], argv[i + 2], n, ms / n), flush=True)
            W.ibm.close()
            W.lattice.close()
            i += 4
        elif cmd == "ensemble":
            # `ensemble N S [EO OO EN ON]`: the bench's ensemble of N shear
            # boxes, batched (collide orders EO / OO and nontemporal levels
            # EN / ON for even / odd steps, default 0), four settle steps,
            # then S timed steps (two per graph launch)
            from . import workloads
            from .ensemble import Ensemble
            n_boxes, n = int(argv[i + 1]), int(argv[i + 2])
            ws = workloads.shear_ensemble("f64", n_boxes=n_boxes)
            ens = Ensemble([(w.lattice, w.ibm) for w in ws])
            shape = [int(a) for a in argv[i + 3:i + 7]] if len(argv) >= i + 7 and argv[i + 3].isdigit() else None
            if shape:
                ens.set_launch(0, shape[0], shape[2])
                ens.set_launch(1, shape[1], shape[3])
            ens.steps(4)
            capi.call("bl_sync")
            ms = ens.time_steps(n)
            print("ensemble %d boxes: batched %s, %.4f ms per step" % (n_boxes, ens.batched, ms / n), flush=True)
            ens.close()
            for w in ws:
                w.ibm.close()
                w.lattice.close()
            i += 7 if shape else 3
        elif cmd == "copy":
            import ctypes
            nbytes, n = int(argv[i + 1]), int(argv[i + 2])
            ms = ctypes.c_double()
            capi.call("bl_device_copy", nbytes, n, ctypes.byref(ms))
            print("copy %d bytes: %.4f ms each, %.1f GB/s (read + write)" %
                  (nbytes, ms.value / n, 2 * nbytes / (ms.value / n * 1e-3) / 1e9), flush=True)
            i += 3
        else:
            raise SystemExit("unknown command %r" % cmd)


def _csv_rows(outdir, suffix):
    rows = []
    for f in sorted(glob.glob(os.path.join(outdir, "**", "*" + suffix), recursive=True)):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def stats(outdir, top=25):
    """Kernel table from a --kernel-trace --stats run (kernel_stats.csv)."""
    rows = _csv_rows(outdir, "kernel_stats.csv")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | total us | avg us | min us | max us | % |")
    print("|---|---|---|---|---|---|---|")
    for r in rows[:top]:
        print("| %s | %s | %.1f | %.3f | %.3f | %.3f | %.2f |" % (
            r["Name"][:90].replace("|", "/"), r["Calls"], float(r["TotalDurationNs"]) / 1e3,
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))


def _counter_per_dispatch(outdir, counter, kernel_substr):
    vals = {}
    for r in _csv_rows(outdir, "counter_collection.csv"):
        if r["Counter_Name"] != counter or kernel_substr not in r["Kernel_Name"]:
            continue
        key = r["Dispatch_Id"]
        vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
    return list(vals.values())


def traffic(fetch_dir, write_dir, copy_bytes, kernel_substr):
    """HBM bytes per launch of `kernel_substr` from FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled by what the same counter reads for the
    plain copy kernel of known size.  Several kernels, comma-separated (the
    two AA parities' kernels, launched equally often): the mean of their
    medians."""
    cb = float(copy_bytes)
    out = {}
    subs = kernel_substr.split(",")
    for name, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        copy = _counter_per_dispatch(d, name, "k_copy16")
        if not copy:
            raise SystemExit("no %s samples for k_copy16 in %s" % (name, d))
        meds = []
        for sub in subs:
            kern = _counter_per_dispatch(d, name, sub)
            if not kern:
                raise SystemExit("no %s samples for %s in %s" % (name, sub, d))
            meds.append(sorted(kern)[len(kern) // 2])
            if len(subs) > 1:
                print("%s: %s %.1f KiB/dispatch (median)" % (name, sub, meds[-1]))
        copy_kib = sorted(copy)[len(copy) // 2]
        factor = cb / (copy_kib * 1024.0)
        kern_kib = sum(meds) / len(meds)
        out[name] = (kern_kib, copy_kib, factor, kern_kib * 1024.0 * factor)
    for name, (k, c, f, b) in out.items():
        print("%s: kernel %.1f KiB/dispatch (median of dispatches%s), copy kernel %.1f KiB for %d bytes, "
              "factor %.4f -> %.4e bytes" % (name, k, ", mean over the kernels" if len(subs) > 1 else "", c, int(cb),
                                             f, b))
    total = out["FETCH_SIZE"][3] + out["WRITE_SIZE"][3]
    print("traffic_bytes_per_launch: %.6e" % total)
    return total


def traffic_table(fetch_dir, write_dir, copy_bytes, trace_dir, top=20):
    """Per-kernel HBM bytes per dispatch (FETCH_SIZE, WRITE_SIZE scaled as
    in `traffic`, medians over dispatches) with the average duration of the
    same kernel from a --kernel-trace --stats run of the same command, and
    the resulting bandwidth against 8 TB/s."""
    cb = float(copy_bytes)
    per = {}
    for name, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _csv_rows(d, "counter_collection.csv")
        disp = {}
        for r in rows:
            if r["Counter_Name"] != name:
                continue
            key = (r["Kernel_Name"], r["Dispatch_Id"])
            disp[key] = disp.get(key, 0.0) + float(r["Counter_Value"])
        byk = {}
        for (k, _), v in disp.items():
            byk.setdefault(k, []).append(v)
        copy = [v for k, vs in byk.items() if "k_copy16" in k for v in vs]
        if not copy:
            raise SystemExit("no %s samples for k_copy16 in %s" % (name, d))
        factor = cb / (sorted(copy)[len(copy) // 2] * 1024.0)
        for k, vs in byk.items():
            if "k_copy16" in k:
                continue
            per.setdefault(k, {})[name] = sorted(vs)[len(vs) // 2] * 1024.0 * factor
    dur = {r["Name"]: float(r["AverageNs"]) for r in _csv_rows(trace_dir, "kernel_stats.csv")}
    rows = []
    for k, v in per.items():
        f, w = v.get("FETCH_SIZE", 0.0), v.get("WRITE_SIZE", 0.0)
        ns = dur.get(k)
        gbs = (f + w) / ns if ns else float("nan")
        rows.append((k, f, w, ns, gbs))
    rows.sort(key=lambda r: -(r[3] or 0))
    print("| kernel | fetch MB | write MB | avg us | GB/s | % of 8 TB/s |")
    print("|---|---|---|---|---|---|")
    for k, f, w, ns, gbs in rows[:top]:
        print("| %s | %.2f | %.2f | %s | %.0f | %.1f |" % (
            k[:90].replace("|", "/"), f / 1e6, w / 1e6, "%.2f" % (ns / 1e3) if ns else "-", gbs, 100 * gbs / 8000.0))


def main(argv):
    if argv and argv[0] == "stats":
        stats(argv[1], int(argv[2]) if len(argv) > 2 else 25)
    elif argv and argv[0] == "traffic":
        traffic(argv[1], argv[2], argv[3], argv[4])
    elif argv and argv[0] == "traffic_table":
        traffic_table(argv[1], argv[2], argv[3], argv[4], int(argv[5]) if len(argv) > 5 else 20)
    else:
        run(argv)


if __name__ == "__main__":
    main(sys.argv[1:])
