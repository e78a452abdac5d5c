"""Prints the per-kernel summary (calls, total and average duration) of a rocprofv3 SQLite
result database:  python profiles/summarize_rocpd.py <results.db>"""
import sqlite3
import sys

con = sqlite3.connect(sys.argv[1])
rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
print(f"{'kernel':<70} {'calls':>6} {'total_us':>11} {'avg_us':>9} {'pct':>6}")
for name, calls, tot, avg, pct in rows:
    print(f"{name[:70]:<70} {calls:>6} {tot:>11.1f} {avg:>9.2f} {pct:>6.1f}")
