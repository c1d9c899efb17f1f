"""Mirror of the reference package ``splinetracers`` (only the fast ellipsoid tracer)."""
