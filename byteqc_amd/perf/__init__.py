"""Workload definitions and the algorithmic work model used by bench.py."""
