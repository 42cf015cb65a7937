"""The viewer's use of the path (view/ of the reference): the FP16 frame cache and the boundary layers a live
run reads per frame. The window, navigation and drawing are not part of this package."""
