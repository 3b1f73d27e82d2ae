"""Copies the reference's stored GA search runs, as data, into
tests/golden/gasearch/: for each instance whose evaluation cache was kept
(results/ga40/logs/<inst>.evcache: every proxy and confirmation value the
search computed, hexfloats) its search log (results/ga40/logs/
<inst>_search.log: per-generation best, mean, distinct genomes and
evaluation count, the finalists and the champion), gzipped.

Run from the repository root where the reference is present:
    python tests/golden/copy_ga_search_logs.py
"""
import gzip
import os
import shutil

SRC = "/root/reference/results/ga40/logs"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gasearch")
INSTANCES = ["c204", "c205", "pr14", "pr15", "rc202", "rc205"]


def main() -> None:
    os.makedirs(DST, exist_ok=True)
    for inst in INSTANCES:
        for name in (inst + ".evcache", inst + "_search.log"):
            with open(os.path.join(SRC, name), "rb") as fi, gzip.open(os.path.join(DST, name + ".gz"), "wb") as fo:
                shutil.copyfileobj(fi, fo)


if __name__ == "__main__":
    main()
