"""The LAMMPS particle dump of the reference's OvitoOutput restated in
Python from Output.jl:146-212 (plotParticles_2D) and :214-288
(plotParticles_3D), as the checker of the library's dump writer.  Numbers go
through Julia's print(Float64) rule (jump_amd.fix.julia_float)."""
from jump_amd.fix import julia_float as jf


def dump_text(dim, lims, solids, outs):
    """solids: list of dicts with pos (n,d), vel (n,d), stress (n,d,d) full
    tensors, alpha (n), vm (n), elastic (bool)."""
    n = sum(len(s["pos"]) for s in solids)
    out = ["ITEM: TIMESTEP \n", "0\n", "ITEM: NUMBER OF ATOMS\n", f"{n} \n", "ITEM: BOX BOUNDS sm sm sm\n",
           f"0 {jf(lims[0])}\n", f"0 {jf(lims[1])}\n", f"0 {jf(lims[2])}\n" if dim == 3 else "0 0\n",
           "ITEM: ATOMS id type x y z "]
    for v in outs:
        out.append(v)
        out.append(" ")
    out.append("\n")
    pid = 0
    for s, sol in enumerate(solids, start=1):
        x, vel, st = sol["pos"], sol["vel"], sol["stress"]
        for p in range(len(x)):
            pid += 1
            if dim == 2:
                out.append(f"{pid} {s} {jf(x[p][0])} {jf(x[p][1])} 0 ")
            else:
                out.append(f"{pid} {s} {jf(x[p][0])} {jf(x[p][1])} {jf(x[p][2])} ")
            for v in outs:
                if v == "pstrain":
                    out.append("0 " if sol["elastic"] else f"{jf(sol['alpha'][p])} ")
                elif v == "vonMises":
                    out.append("0 " if sol["elastic"] else f"{jf(sol['vm'][p])} ")
                elif v == "sigmaxx" and dim == 2:
                    out.append(f"{jf(st[p][0][0])} ")
                elif v == "sigmayy" and dim == 2:
                    out.append(f"{jf(st[p][1][1])} ")
                elif v == "pressure":
                    if dim == 2:
                        out.append(f"{jf(st[p][0][0] + st[p][1][1])} ")
                    else:
                        out.append(f"{jf(st[p][0][0] + st[p][1][1] + st[p][2][2])}  ")
                elif v == "vx":
                    out.append(f"{jf(vel[p][0])}  ")
                elif v == "vy":
                    out.append(f"{jf(vel[p][1])}  ")
                elif v == "vz" and dim == 3:
                    out.append(f"{jf(vel[p][2])}  ")
            out.append("\n")
    return "".join(out)
