"""Operator classes with the reference's names, signatures and layouts (processing/*.py)."""
