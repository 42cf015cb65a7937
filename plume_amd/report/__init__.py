"""Reports measured from the sheet frames (report/ of the reference): the numbers file."""
