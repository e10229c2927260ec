"""Viewer for the CSV trajectories written by nbody_cli (one row per step:
t, x0, y0, z0, x1, y1, z1, ...; the row format of the reference's main.cpp:
77-95).  Python 3 counterpart of the reference's script/animation.py: an
animated 3-D scatter of the particles.

    python -m nbody_amd.viewer particles.csv              # interactive window
    python -m nbody_amd.viewer particles.csv --save a.gif  # write an animation
    python -m nbody_amd.viewer particles.csv --frame -1 --save last.png
"""
from __future__ import annotations

import argparse
import csv

import numpy as np


def read_csv(path):
    """(times[T], positions[T, N, 3]) from a trajectory file."""
    times, rows = [], []
    with open(path, newline="") as f:
        for row in csv.reader(f):
            if not row:
                continue
            vals = [float(v) for v in row]
            if (len(vals) - 1) % 3:
                raise ValueError(f"row of {len(vals)} values is not t followed by x,y,z triples")
            times.append(vals[0])
            rows.append(vals[1:])
    if not rows:
        raise ValueError(f"{path} holds no rows")
    n = len(rows[0]) // 3
    if any(len(r) != 3 * n for r in rows):
        raise ValueError("rows hold different particle counts")
    return np.array(times), np.array(rows).reshape(len(rows), n, 3)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("csv")
    ap.add_argument("--save", help="write the animation (.gif/.mp4) or one frame (.png) instead of showing it")
    ap.add_argument("--frame", type=int, help="only this frame (index, negative from the end)")
    ap.add_argument("--fps", type=int, default=25)
    ap.add_argument("--limits", type=float, nargs=2, default=None,
                    help="axis range (default: the trajectory's bounding cube)")
    args = ap.parse_args(argv)
    times, P = read_csv(args.csv)

    import matplotlib
    if args.save:
        matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from matplotlib import animation

    lo, hi = (args.limits if args.limits else (float(P.min()), float(P.max())))
    fig = plt.figure(figsize=(6, 6))
    ax = fig.add_subplot(projection="3d")
    for setlim, lab in ((ax.set_xlim3d, "X"), (ax.set_ylim3d, "Y"), (ax.set_zlim3d, "Z")):
        setlim([lo, hi])
    ax.set_xlabel("X"); ax.set_ylabel("Y"); ax.set_zlabel("Z")
    first = P[args.frame] if args.frame is not None else P[0]
    pts = ax.scatter(first[:, 0], first[:, 1], first[:, 2], s=2)
    title = ax.set_title(f"t = {times[args.frame if args.frame is not None else 0]:.4g}")

    if args.frame is not None:
        if args.save:
            fig.savefig(args.save)
        else:
            plt.show()
        return 0

    def update(i):
        pts._offsets3d = (P[i, :, 0], P[i, :, 1], P[i, :, 2])
        title.set_text(f"t = {times[i]:.4g}")
        return pts, title

    anim = animation.FuncAnimation(fig, update, frames=len(P), interval=1000 / args.fps, blit=False)
    if args.save:
        writer = "pillow" if args.save.endswith(".gif") else None
        anim.save(args.save, writer=writer, fps=args.fps)
    else:
        plt.show()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
