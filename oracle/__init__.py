"""CPU oracle for the SL-BDF2 step (test infrastructure only; never imported by the product)."""
