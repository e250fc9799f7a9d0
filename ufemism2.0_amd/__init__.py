"""MI355X engine for UFEMISM's SSA/DIVA velocity solve and vertical heat columns."""
