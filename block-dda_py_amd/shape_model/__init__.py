"""Drop-in module of the MI355X block-DDA solver (see bdda_amd)."""
