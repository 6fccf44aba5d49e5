"""Helpers writing the reference's input files (41-tag parameter file,
src/read_params.c:65-227; 6-column ASCII catalogue, src/read_data.c:64)."""
import numpy as np

TAGS = [
    ("NumFilesInParallel", 1), ("NX", 32), ("NY", 32), ("NZ", 32), ("XMIN", 0.0),
    ("XMAX", 1000.0), ("YMIN", 0.0), ("YMAX", 1000.0), ("ZMIN", 0.0), ("ZMAX", 1000.0),
    ("DoInterlacing", 0), ("InterpOrder", 2), ("NK", 10), ("NMU", 4), ("Output2D", 0),
    ("OutputLog", 0), ("Mink", 0.0), ("Maxk", 0.1), ("Omega_m", 0.3089), ("Redshift", 0.1),
    ("Periodic", 1), ("LOS", 0), ("Survey", 0), ("Momentum", 0), ("X_Origin", 0.0),
    ("Y_Origin", 0.0), ("Z_Origin", 0.0), ("REDMININ", 0.0), ("REDMAXIN", 1.0),
    ("NOBJ_Max", 1000), ("Coord_Type", 0), ("x_Column", 1), ("y_Column", 2), ("z_Column", 3),
    ("NBAR_Column", -1), ("SkyArea", 1.0), ("FKP_Column", -1), ("FKP_Pk", 0.0),
    ("FKP_Pk_mom", 0.0), ("FKP_Pk_cross", 0.0), ("Odd_Multipoles", 0),
]


def param_text(drop=(), extra=(), **over):
    lines = ["% test parameter file", ""]
    for k, v in TAGS:
        if k in drop:
            continue
        lines.append("%-20s %s    %% comment" % (k, over.get(k, v)))
    for k, v in extra:
        lines.append("%s %s" % (k, v))
    return "\n".join(lines) + "\n"


def write_params(path, **kw):
    with open(path, "w") as f:
        f.write(param_text(**kw))
    return path


def write_catalogue(path, pos, vel, header=True):
    with open(path, "w") as f:
        if header:
            f.write("# x y z vx vy vz\n")
        for p, v in zip(pos, vel):
            f.write("%.17g %.17g %.17g %.17g %.17g %.17g\n" % (*p, *v))
    return path
