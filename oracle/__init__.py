"""Test-only oracle of the reference algorithm (see gpe_oracle.py).  Not product code."""
