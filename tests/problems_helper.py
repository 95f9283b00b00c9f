"""Small helpers shared by tests."""
import importlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

P = importlib.import_module("hpc-reading-course_amd.problems")


def pmin(order):
    return P.min_points(order)
