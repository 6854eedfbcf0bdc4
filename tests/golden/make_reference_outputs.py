"""Extract small fixtures from the solution files the reference ships.

The reference keeps CGNS outputs of its own runs under zone_FD/cases and
CongFD1.0/cases, with their per-step logs.  This script reads them with the
package's CGNS reader (cfd_inwsl_amd/cgns.py) and writes:

* reference_outputs.npz: cell-centred Density / XVelocity / Pressure of the
  1D end states listed in CASES (TENO.cgns, TENOS.cgns, TENOAS.cgns and,
  where present, TENOLAD.cgns), the t=0 and t=0.2 states and the logged dt
  sequence of the "Sod - congTCNS5CT5 - 101" run, the t=0.1 and t=0.2 states
  of the "Sod - whfzycTCNSNmyASF002_1 - 101" run, and the 512x128 double-Mach
  density at t=0.2 (float32: it is compared loosely);
* reference_outputs.json: the run each array comes from (the reference file,
  the driver's case number and info.txt values), and SHA-256 digests of the
  fields and coordinates of initial-state files, which the driver's
  --initial-only output must reproduce bit for bit;
* CongFD1.0's runs: the two 200-cell Lax series (every output's digests, the
  last output's fields) and the two 64x256 Rayleigh-Taylor runs of its
  gamma = 5/3 build (digests of the t = 0 and t = 1.95 fields, the t = 1.95
  density as float32).

Run in a checkout that has the reference tree (default /root/reference):
    python tests/golden/make_reference_outputs.py [REFERENCE_ROOT]
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from cfd_inwsl_amd import cgns  # noqa: E402

# 1D end states (TENO5 = Teno5_Z, info.txt inter 11); zone and iMax from the
# files' own coordinates, outputDt the value the reference's run used (the
# dt clamp at output times changes the step sequence).
CASES = {
    "sod": dict(file="zone_FD/cases/sod tube    t=0.2000/TENO.cgns", case=0, t_end=0.2, output_dt=0.01),
    "shu_osher": dict(file="zone_FD/cases/Shu-Osher  t=1.8000/200/TENO.cgns", case=1, t_end=1.8, output_dt=1.8),
    "lax": dict(file="zone_FD/cases/lax sod tube    t=0.1400/TENO.cgns", case=2, t_end=0.14, output_dt=0.01),
    "double_rarefaction": dict(file="zone_FD/cases/ShuangXiShuBo   t=1.0000/TENO.cgns", case=5, t_end=1.0,
                               output_dt=0.01),
    # reflecting walls (SYMMETRY1D); that run used CFL 0.1 (0.5 blows up, in
    # the reference as here); its TENOS.cgns came from other settings
    "woodward_colella": dict(file="zone_FD/cases/Woodward-Colella   t=0.0380/TENO.cgns", case=4, t_end=0.038,
                             output_dt=0.038, cfl=0.1, tenos_pinned=False),
}
INITIAL = {
    "sod_101": dict(file="zone_FD/cases/Sod - whfzycTCNSNmyASF002_1 - 101 - t=0.0000.cgns", dim=1, case=0),
    "sod_10000": dict(file="zone_FD/cases/EULERWCNS5t=0.0000.cgns", dim=1, case=0),
    "lax_200": dict(file="CongFD1.0/cases/EULERWCNS5t=0.0000.cgns", dim=1, case=2),
    "woodward_colella_40": dict(file="zone_FD/cases/Woodward_Colella - whfCOMPARE - 41 - t=0.0000.cgns", dim=1,
                                case=4),
    "woodward_colella_400": dict(file="zone_FD/cases/Woodward_Colella - whfCOMPARE - 401 - t=0.0000.cgns", dim=1,
                                 case=4),
}
SOD_LOG = "zone_FD/cases/Sod - congTCNS5CT5 - 101"
# whf_zyc_TcnsN_myASF002_1 (InterMethod whfzycTCNSNmyASF002_1 = 33), outputs at t = 0.1 and 0.2
SOD_ZYC = "zone_FD/cases/Sod - whfzycTCNSNmyASF002_1 - 101"
DMR = "zone_FD/cases/Double Mach    t=0.2000/512, 128 - TENO   t=0.2000.cgns"
# CongFD1.0's own runs (its Euler path: Roe flux hard-wired, no 1D limiter,
# three-division MND6): two 200-cell Lax series with outputs every 0.01 up to
# 0.14, and two 64x256 Rayleigh-Taylor runs of the gamma = 5/3 build at t = 1.95
C1_LAX = {
    "c1_lax_f203": "CongFD1.0/cases/Lax - TENO-AS-myF203_NoS - 201 - t={t}.cgns",
    "c1_lax_tenos": "CongFD1.0/cases/EULERWCNS5t={t}.cgns",
}
C1_RTI = {
    "c1_rti_whftcnsa": "CongFD1.0/cases/RTI - WHFTCNSA - 65x257 - t={t}.cgns",
    "c1_rti_whftcnsaf002": "CongFD1.0/cases/RTI - WHFTCNSAF002 - 65x257 - t={t}.cgns",
}
FIELDS = ("Density", "XVelocity", "YVelocity", "ZVelocity", "Pressure")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def zone_of(sol):
    x = sol["CoordinateX"]
    xs = x if x.ndim == 1 else x[:, 0]
    imax = [len(xs), 2, 2]
    zone = [float(xs[0]), float(xs[-1]), 0.0, 1.0, 0.0, 0.0]
    if "CoordinateY" in sol:
        y = sol["CoordinateY"][0, :]
        imax[1] = len(y)
        zone[2:4] = [float(y[0]), float(y[-1])]
    return zone, imax


def tree_listing(node, path=""):
    out = []
    for c in node.children:
        p = path + "/" + c.name
        out.append([p, c.label, c.type, list(c.data.shape) if c.data is not None else None])
        out += tree_listing(c, p)
    return out


def main(root):
    arrays, meta = {}, {"cases": {}, "initial": {}}
    for key, c in CASES.items():
        sol = cgns.read_solution(os.path.join(root, c["file"]))
        zone, imax = zone_of(sol)
        for f in ("Density", "XVelocity", "Pressure"):
            arrays[f"{key}/{f}"] = sol[f]
        meta["cases"][key] = dict(dict(cfl=0.5, tenos_pinned=True), **c, dim=1, zone=zone, imax=imax, inter=11,
                                  flux="hllc")
        # the same run with TENO-S (Teno5_CongZ, InterMethod WCNS5CONGZ = 8)
        sol = cgns.read_solution(os.path.join(root, c["file"].replace("TENO.cgns", "TENOS.cgns")))
        for f in ("Density", "XVelocity", "Pressure"):
            arrays[f"{key}_tenos/{f}"] = sol[f]
        # TENOAS.cgns (whf_ai_TcnsN_AS_1/_2, InterMethod 26/27) and, where the
        # reference has it, TENOLAD.cgns (whf_TcnsN_LAD, InterMethod 21)
        for variant in ("TENOAS", "TENOLAD"):
            path = os.path.join(root, c["file"].replace("TENO.cgns", variant + ".cgns"))
            if os.path.exists(path):
                sol = cgns.read_solution(path)
                for f in ("Density", "XVelocity", "Pressure"):
                    arrays[f"{key}_{variant.lower()}/{f}"] = sol[f]
    for key, c in INITIAL.items():
        sol = cgns.read_solution(os.path.join(root, c["file"]))
        zone, imax = zone_of(sol)
        meta["initial"][key] = dict(c, zone=zone, imax=imax,
                                    sha256={k: digest(v) for k, v in sol.items()
                                            if k in FIELDS or k.startswith("Coordinate")})
    # the logged Sod run: per-step "time = .. dt=..  rhoLinf = .." (two runs appended)
    with open(os.path.join(root, SOD_LOG + ".txt")) as f:
        log = [[float(v) for v in re.findall(r"=\s*([-+\d.eE]+)", line)] for line in f if line.strip()]
    log = np.array(log)
    restart = np.nonzero(np.diff(log[:, 0]) < 0)[0]
    first = log[: restart[0] + 1] if len(restart) else log
    arrays["sod_ct5/log_time_dt_rholinf"] = first
    for t in ("0.0000", "0.2000"):
        sol = cgns.read_solution(os.path.join(root, f"{SOD_LOG} - t={t}.cgns"))
        for f in ("Density", "XVelocity", "Pressure"):
            arrays[f"sod_ct5/t{t}/{f}"] = sol[f]
    zone, imax = zone_of(sol)
    meta["sod_ct5"] = dict(file=SOD_LOG, case=0, dim=1, zone=zone, imax=imax, inter=11, flux="hllc", cfl=0.5,
                           output_dt=0.1, t_end=0.2)
    for t in ("0.1000", "0.2000"):
        sol = cgns.read_solution(os.path.join(root, f"{SOD_ZYC} - t={t}.cgns"))
        for f in ("Density", "XVelocity", "Pressure"):
            arrays[f"sod_zyc/t{t}/{f}"] = sol[f]
    zone, imax = zone_of(sol)
    meta["sod_zyc"] = dict(file=SOD_ZYC, case=0, dim=1, zone=zone, imax=imax, inter=33, flux="hllc", cfl=0.5,
                           output_dt=0.1, t_end=0.2)
    sol = cgns.read_solution(os.path.join(root, DMR))
    zone, imax = zone_of(sol)
    arrays["dmr/Density_f32"] = sol["Density"].astype(np.float32)
    # the double-Mach run wrote an output every 0.01 (zoneMain's Double Mach
    # preset: endStep 20, outputDt 0.01); digests of the four schemes' t = 0.2 fields
    meta["dmr"] = dict(file=DMR, case=4, dim=2, zone=zone, imax=imax, inter=11, flux="hllc", cfl=0.5, t_end=0.2,
                       output_dt=0.01, sha256={})
    for variant in ("TENO", "TENOS", "TENOA", "TENOAS"):
        sol = cgns.read_solution(os.path.join(root, DMR.replace("- TENO ", f"- {variant} ")))
        meta["dmr"]["sha256"][variant] = {f: digest(sol[f].T.reshape(-1))
                                          for f in ("Density", "XVelocity", "YVelocity", "Pressure")}
    # CongFD1.0 Lax series: the t = 0.14 fields in full, SHA-256 digests of
    # every output's fields (t = 0.01 .. 0.14)
    for key, pattern in C1_LAX.items():
        meta[key] = dict(file=pattern, case=2, dim=1, cfl=0.5, output_dt=0.01, n_out=14, sha256=[])
        for k in range(1, 15):
            sol = cgns.read_solution(os.path.join(root, pattern.format(t=f"{0.01 * k:.4f}")))
            meta[key]["sha256"].append({f: digest(sol[f]) for f in ("Density", "XVelocity", "Pressure")})
        zone, imax = zone_of(sol)
        meta[key].update(zone=zone, imax=imax)
        for f in ("Density", "XVelocity", "Pressure"):
            arrays[f"{key}/t0.1400/{f}"] = sol[f]
    # CongFD1.0 Rayleigh-Taylor runs: digests of the t = 0 and t = 1.95 fields
    # (cells x fastest), the t = 1.95 density as float32 for error reports
    for key, pattern in C1_RTI.items():
        meta[key] = dict(file=pattern, case=3, dim=2, cfl=0.5, output_dt=1.95)
        for t in ("0.0000", "1.9500"):
            sol = cgns.read_solution(os.path.join(root, pattern.format(t=t)))
            meta[key][f"sha256_t{t}"] = {f: digest(sol[f].T.reshape(-1))
                                        for f in ("Density", "XVelocity", "YVelocity", "Pressure")}
        zone, imax = zone_of(sol)
        meta[key].update(zone=zone, imax=imax)
        arrays[f"{key}/t1.9500/Density_f32"] = sol["Density"].T.reshape(-1).astype(np.float32)
    tree = cgns.read_tree(os.path.join(root, CASES["sod"]["file"]))
    meta["tree_1d"] = tree_listing(tree)
    meta["root_attrs"] = {k: v for k, v in tree.attrs.items() if isinstance(v, str)}
    np.savez_compressed(os.path.join(HERE, "reference_outputs.npz"), **arrays)
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
