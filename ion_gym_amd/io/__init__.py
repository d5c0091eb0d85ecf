"""On-disk formats of solved fields: content-keyed array entries
(fa_cache) and the per-electrode basis layouts on top of them
(basis_cache)."""
